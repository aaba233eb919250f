"""The allocation-free exact product (amm_exact_product_noalloc) against the
context entry point amm_exact_product on the same inputs: two calls in a
row, the second on a smaller matrix, each within 1e-5 relative of the
context's result and of the FP64 product."""
import numpy as np
import pytest

from helpers import bf16_bits, bf16_to_f32, lowrank_plus_noise, rel, to_cols

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(A, ld, bf16):
    cols = to_cols(A, ld)
    if bf16:
        bits = bf16_bits(cols)
        return torch.from_numpy(bits.view(np.int16)).cuda().view(torch.bfloat16), bf16_to_f32(bits)[:, :A.shape[0]].T
    return torch.from_numpy(cols).cuda(), A


def test_noalloc_exact_product_matches_context_entry_point():
    from eigenvaluecomputation_by_amm_amd import AmmContext
    from eigenvaluecomputation_by_amm_amd.api import exact_product_noalloc
    cases = [(16384, 4096, None, False), (3000, 5000, 3004, False), (1000, 700, 1008, True)]
    outs = []
    for k, (d, n, ld, bf16) in enumerate(cases):
        print(f"\n[gpu-case] noalloc exact product d={d} n={n} ld={ld} bf16={bf16}", flush=True)
        A = lowrank_plus_noise(d, n, seed=40 + k) + 0.1
        t, Ah = _dev(A, ld, bf16)
        x = torch.from_numpy(np.random.default_rng(50 + k).standard_normal(d).astype(np.float32)).cuda()
        y = exact_product_noalloc(t, x, d=d)
        y2 = exact_product_noalloc(t, x, d=d)  # the second call reuses the same scratch
        torch.cuda.synchronize()
        with AmmContext(t, d=d) as ctx:
            yc = ctx.exact_product(x).cpu().numpy()
        yn = y.cpu().numpy()
        A64 = np.asarray(Ah, np.float64)
        ref = A64 @ (A64.T @ x.cpu().numpy().astype(np.float64))
        assert rel(yn, yc) < 1e-5
        assert rel(yn, ref) < 1e-5
        assert np.array_equal(yn, y2.cpu().numpy())
        outs.append(yn)
    assert [o.size for o in outs] == [c[0] for c in cases]


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from eigenvaluecomputation_by_amm_amd import AmmError
    from eigenvaluecomputation_by_amm_amd.api import exact_product_noalloc
    # 2^21 rows: x, y and one partial of A t alone exceed 16 MiB
    d, n = 1 << 21, 4
    t = torch.zeros((n, d), dtype=torch.float32, device="cuda")
    x = torch.zeros(d, dtype=torch.float32, device="cuda")
    with pytest.raises(AmmError, match="scratch"):
        exact_product_noalloc(t, x)
