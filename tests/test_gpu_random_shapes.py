"""Seeded random shapes through the whole single-matrix path: ragged d, n
and leading dimensions, FP32 and BF16 storage, all four kernels with and
without variance reduction, c from 1 to n.  Each case runs the power method
on the device and the oracle on the same stored values and the device's own
column weights (so the samples are the same): identical step kinds and
Rayleigh quotients within 1e-4; plus the exact product within 1e-5 of FP64."""
import numpy as np
import pytest

from helpers import bf16_bits, bf16_to_f32, heavy_columns, rel, to_cols
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KERNELS = ["exact", "previousAMM", "proposedAMM", "proposedAMM_v2"]


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    bf16 = bool(rng.integers(2))
    d = int(rng.integers(1, 6000))
    n = int(rng.integers(1, 2500))
    width = 8 if bf16 else 4
    ld = (d + width - 1) // width * width + width * int(rng.integers(0, 3))
    kernel = KERNELS[seed % 4]
    c = int(rng.integers(1, n + 1))
    vr = int(rng.integers(2, 6)) if kernel == "proposedAMM_v2" else int(rng.integers(0, 5))
    return d, n, ld, bf16, kernel, c, vr


@pytest.mark.parametrize("seed", range(16))
def test_random_shape_run_and_product(seed):
    d, n, ld, bf16, kernel, c, vr = _case(seed)
    print(f"\n[gpu-case] random seed={seed} d={d} n={n} ld={ld} bf16={bf16} {kernel} c={c} vr={vr}",
          flush=True)
    from eigenvaluecomputation_by_amm_amd import AmmContext, RunParams
    A = heavy_columns(d, n, seed=seed, rank=min(4, d, n)) + 0.1
    cols = to_cols(A, ld)
    if bf16:
        bits = bf16_bits(cols)
        t = torch.from_numpy(bits.view(np.int16)).cuda().view(torch.bfloat16)
        vals = bf16_to_f32(bits)
    else:
        t = torch.from_numpy(cols).cuda()
        vals = cols
    run = dict(kernel=kernel, step=8, vr_period=vr)
    if kernel != "exact":
        run["c"] = c
    x = np.random.default_rng(seed).standard_normal(d).astype(np.float32)
    with AmmContext(t, d=d) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=seed + 5, **run))
        modes = ctx.last_modes(run["step"])
        y = ctx.exact_product(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = O.power_method(vals, d, seed=seed + 5, w_plain=nrm2, w_cent=cnrm2, **run)
    dm = modes.copy()
    dm[dm == 4] = 1  # exact without memo reports as exact
    assert np.array_equal(dm, ref["modes"].astype(np.uint32))
    assert np.max(np.abs(res.eig - ref["eig"]) / np.maximum(np.abs(ref["eig"]), 1e-30)) < 1e-4
    A64 = vals[:, :d].T.astype(np.float64)
    assert rel(y, A64 @ (A64.T @ x)) < 1e-5
