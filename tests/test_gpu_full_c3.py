"""The image-like BF16 configuration at full size: 40000 Gaussian-random-
field images of 64 x 64 pixels (SE kernel, length scale 4 px, min-max to
[0, 1], 5% uniform noise, rows centred), d = 4096, n = 40000, stored in
BF16.  proposedAMM with c = 2048 (mean-shifted sampling,
powerMethod_by_proposed.cuh:105-235) on the batched MFMA engine, 16 trials:

* every trial ends within 1e-5 of the FP64 eigvalsh of the stored BF16
  matrix's Gram and reaches 1e-3;
* trials 0 and 1 follow the single-matrix run with the same trial word
  (same samples): identical step kinds, Rayleigh quotients within 1e-4;
* the FP32 exact power method on the FP32 data ends within 1e-6 of the
  FP64 eigvalsh of the FP32 matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RUN = dict(kernel="proposedAMM", c=2048, vr_period=5, step=30)


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def relmax(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def lam64(M):
    M64 = M.double()
    G = M64.T @ M64
    del M64
    return float(torch.linalg.eigvalsh(G)[-1])


@pytest.fixture(scope="module")
def c3():
    from eigenvaluecomputation_by_amm_amd.workloads import grf_images
    banner("C3 generate 40000 GRF images 64x64 + FP64 eigvalsh")
    A32 = grf_images(side=64, n=40000)
    Ab = A32.to(torch.bfloat16)
    yield dict(A32=A32, Ab=Ab, lam_b=lam64(Ab), lam_32=lam64(A32))
    del A32, Ab
    torch.cuda.empty_cache()


@pytest.mark.timeout(600)
def test_c3_batched_bf16_against_eigvalsh(c3):
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    banner("C3 batched proposedAMM c=2048 x16 BF16")
    Ab, lam = c3["Ab"], c3["lam_b"]
    params = RunParams(seed=3, trial=0, **RUN)
    with AmmBatch([dict(A=Ab, c=RUN["c"], seed=3)], trials=16) as b:
        r = b.power_method(params)
        modes = b.last_modes(RUN["step"])
    err = np.abs(r.eig[0] - lam) / lam  # [trial][step]
    assert err[:, -1].max() < 1e-5, err[:, -1].max()
    assert all((e <= 1e-3).any() for e in err)
    with AmmContext(Ab) as ctx:
        for k in (0, 1):
            s = ctx.power_method(RunParams(seed=3, trial=k, **RUN))
            assert np.array_equal(ctx.last_modes(RUN["step"]), modes[0, k])
            assert relmax(r.eig[0, k], s.eig) < 1e-4, (k, relmax(r.eig[0, k], s.eig))


@pytest.mark.timeout(600)
def test_c3_exact_fp32_against_eigvalsh(c3):
    from eigenvaluecomputation_by_amm_amd import AmmContext, RunParams
    banner("C3 exact FP32")
    lam = c3["lam_32"]
    with AmmContext(c3["A32"]) as ctx:
        r = ctx.power_method(RunParams(kernel="exact", step=60, seed=1))
    assert abs(r.eig[-1] - lam) / lam < 1e-6
    # BF16 storage moves lambda_max by far less than the 1e-3 target
    assert abs(c3["lam_b"] - lam) / lam < 1e-4
