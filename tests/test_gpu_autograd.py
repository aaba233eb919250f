"""Differentiable sigma_max: forward against FP64 SVD, backward against the
analytic gradient of torch.linalg.svdvals (FP64, CPU) and against central
finite differences of the FP64 singular value along 32 directions.

sigma_max(W) is sqrt of the reference's lambda (powerMethod_by_exact.cuh:83);
d sigma / d W = u1 v1^T is the textbook derivative of a simple singular
value, which the FP64 checks pin independently of this library.
"""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def planted(rows, cols, seed, top=10.0, second=7.0):
    """rows x cols FP64 with singular values top, second, then a slow decay:
    the power method on W^T W contracts by (second/top)^2 = 0.49 a step."""
    g = torch.Generator().manual_seed(seed)
    k = min(rows, cols)
    Q1, _ = torch.linalg.qr(torch.randn(rows, k, generator=g, dtype=torch.float64))
    Q2, _ = torch.linalg.qr(torch.randn(cols, k, generator=g, dtype=torch.float64))
    s = torch.cat([torch.tensor([top, second], dtype=torch.float64),
                   second * 0.95 ** torch.arange(1, k - 1, dtype=torch.float64)])
    return (Q1 * s) @ Q2.T


def _mods():
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max
    return RunParams, sigma_max


def _svd_grad(W64):
    Wr = W64.clone().requires_grad_(True)
    torch.linalg.svdvals(Wr)[0].backward()
    return Wr.grad


@pytest.mark.parametrize("rows,cols", [(384, 256), (256, 384), (300, 203), (1000, 64)])
def test_sigma_forward_backward_f32(rows, cols):
    banner(f"sigma_max f32 rows={rows} cols={cols}")
    RunParams, sigma_max = _mods()
    W64 = planted(rows, cols, seed=rows + cols)
    W = W64.float().cuda().requires_grad_(True)
    s = sigma_max(W, RunParams(kernel="exact", step=100))
    s.backward()
    ref_s = torch.linalg.svdvals(W64.float().double())[0].item()
    assert abs(s.item() - ref_s) <= 2e-6 * ref_s
    G = W.grad.double().cpu()
    Gref = _svd_grad(W64.float().double())
    assert W.grad.dtype == torch.float32 and W.grad.shape == W.shape
    assert (G - Gref).abs().max().item() <= 2e-5 * Gref.abs().max().item()


def test_sigma_gradient_finite_differences():
    """<grad, D> against (sigma(W + hD) - sigma(W - hD)) / 2h of the FP64
    singular value, 32 random unit directions plus the gradient itself."""
    banner("sigma_max finite differences 256x192")
    RunParams, sigma_max = _mods()
    W64 = planted(256, 192, seed=7)
    W = W64.float().cuda().requires_grad_(True)
    sigma_max(W, RunParams(kernel="exact", step=100)).backward()
    G = W.grad.double().cpu()
    Wf = W64.float().double()
    h = 1e-3
    g = torch.Generator().manual_seed(11)
    dirs = [torch.randn(256, 192, generator=g, dtype=torch.float64) for _ in range(32)]
    dirs.append(G.clone())
    for D in dirs:
        D = D / D.norm()
        fd = (torch.linalg.svdvals(Wf + h * D)[0] - torch.linalg.svdvals(Wf - h * D)[0]).item() / (2 * h)
        an = (G * D).sum().item()
        assert abs(fd - an) <= 1e-5 + 1e-4 * abs(fd), (fd, an)
    # the library's own forward along the gradient: d sigma = h ||G|| = h
    Gn = (G / G.norm()).float().cuda()
    hs = 0.05
    with torch.no_grad():
        sp = sigma_max(W.detach() + hs * Gn, RunParams(kernel="exact", step=100)).item()
        sm = sigma_max(W.detach() - hs * Gn, RunParams(kernel="exact", step=100)).item()
    assert abs((sp - sm) / (2 * hs) - 1.0) < 5e-4


def test_sigma_amm_forward():
    """AMM forward (previous, c = n / 2, period-5 exact steps) converges to the
    same sigma and gradient within the sampling tolerance of a gapped matrix."""
    banner("sigma_max previousAMM 512x2048")
    RunParams, sigma_max = _mods()
    W64 = planted(2048, 512, seed=3)
    W = W64.float().cuda().requires_grad_(True)
    s = sigma_max(W, RunParams(kernel="previousAMM", c=1024, step=60, vr_period=5, seed=9))
    s.backward()
    ref_s = torch.linalg.svdvals(W64.float().double())[0].item()
    assert abs(s.item() - ref_s) <= 1e-5 * ref_s
    Gref = _svd_grad(W64.float().double())
    assert (W.grad.double().cpu() - Gref).abs().max().item() <= 2e-3 * Gref.abs().max().item()


def test_sigma_bf16():
    banner("sigma_max bf16 384x256")
    RunParams, sigma_max = _mods()
    W64 = planted(384, 256, seed=5)
    W = W64.to(torch.bfloat16).cuda().requires_grad_(True)
    s = sigma_max(W, RunParams(kernel="exact", step=100))
    s.backward()
    Wq = W.detach().double().cpu()  # the BF16-rounded matrix
    ref_s = torch.linalg.svdvals(Wq)[0].item()
    assert abs(s.item() - ref_s) <= 2e-6 * ref_s
    assert W.grad.dtype == torch.bfloat16
    Gref = _svd_grad(Wq)
    assert (W.grad.double().cpu() - Gref).abs().max().item() <= 2 ** -7 * Gref.abs().max().item()


def test_sigma_context_reuse_sees_new_values():
    """A cached context re-reads the matrix: in-place updates change sigma."""
    banner("sigma_max context reuse")
    RunParams, sigma_max = _mods()
    W = planted(256, 128, seed=1).float().cuda()
    p = RunParams(kernel="exact", step=80)
    s1 = sigma_max(W, p).item()
    W.mul_(2.0)
    s2 = sigma_max(W, p).item()
    assert abs(s2 - 2 * s1) <= 1e-5 * s2


@pytest.mark.parametrize("d,n,ld,dtype", [(256, 100, 256, torch.float32), (203, 77, 203, torch.float32),
                                          (203, 77, 208, torch.float32), (256, 50, 256, torch.bfloat16),
                                          (201, 33, 201, torch.bfloat16), (13, 9, 16, torch.bfloat16)])
def test_rank1_update_bitwise(d, n, ld, dtype):
    """out[:, j] = fl(g v_j) u (FP32, one rounding each), stored FP32 or
    BF16 (round to nearest even); padding rows between d and ld untouched."""
    banner(f"rank1 d={d} n={n} ld={ld} {dtype}")
    from eigenvaluecomputation_by_amm_amd.api import rank1_update
    g = torch.Generator().manual_seed(d * n)
    u = torch.randn(d, generator=g).cuda()
    v = torch.randn(n, generator=g).cuda()
    out = torch.full((n, ld), 7.0, dtype=dtype, device="cuda")
    rank1_update(out, 0.37, u, v, d=d, n=n, ld=ld)
    torch.cuda.synchronize()
    s = (torch.tensor(0.37, dtype=torch.float32) * v.cpu())
    ref = (s[:, None] * u.cpu()[None, :]).to(dtype)
    got = out.cpu()
    bits = torch.int16 if dtype == torch.bfloat16 else torch.int32
    assert torch.equal(got[:, :d].contiguous().view(bits), ref.contiguous().view(bits))
    if ld > d:
        assert torch.all(got[:, d:] == 7.0)


@pytest.mark.parametrize("d,n,ld,dtype", [(256, 100, 256, torch.float32), (203, 77, 208, torch.float32),
                                          (201, 33, 201, torch.bfloat16), (64, 40, 64, torch.bfloat16),
                                          (13, 9, 16, torch.bfloat16)])
def test_rank1_update_batched_matches_single(d, n, ld, dtype):
    """One launch over a stack equals the per-matrix calls bit for bit
    (per-matrix g from a device tensor; padding rows untouched)."""
    banner(f"rank1 batched d={d} n={n} ld={ld} {dtype}")
    from eigenvaluecomputation_by_amm_amd.api import rank1_update, rank1_update_batched
    B = 5
    gen = torch.Generator().manual_seed(d + n)
    U = torch.randn(B, d, generator=gen).cuda()
    V = torch.randn(B, n, generator=gen).cuda()
    g = torch.randn(B, generator=gen).cuda()
    out = torch.full((B, n, ld), 3.0, dtype=dtype, device="cuda")
    rank1_update_batched(out, g, U, V, d=d, n=n, ld=ld)
    ref = torch.full((B, n, ld), 3.0, dtype=dtype, device="cuda")
    gh = g.cpu().tolist()
    for k in range(B):
        rank1_update(ref[k], gh[k], U[k], V[k], d=d, n=n, ld=ld)
    torch.cuda.synchronize()
    bits = torch.int16 if dtype == torch.bfloat16 else torch.int32
    assert torch.equal(out.cpu().view(bits), ref.cpu().view(bits))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_sigma_stack(dtype):
    """A stack of planted matrices in one batched run: sigma and the
    gradient of every matrix against FP64 svdvals (of the stored values)."""
    banner(f"sigma_max stack {dtype}")
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max_stack
    Ws = torch.stack([planted(320, 256, seed=40 + k, top=10.0 + k, second=6.0) for k in range(6)])
    W = Ws.to(dtype).cuda().requires_grad_(True)
    s = sigma_max_stack(W, RunParams(kernel="previousAMM", c=128, step=80, vr_period=5, seed=3))
    (s * torch.arange(1, 7, device="cuda", dtype=torch.float32)).sum().backward()
    Wq = W.detach().double().cpu()
    tol_g = 2e-5 if dtype == torch.float32 else 2 ** -7
    for k in range(6):
        ref = torch.linalg.svdvals(Wq[k])[0].item()
        assert abs(s[k].item() - ref) <= 5e-6 * ref, (k, s[k].item(), ref)
        Gref = (k + 1) * _svd_grad(Wq[k])
        G = W.grad[k].double().cpu()
        assert (G - Gref).abs().max().item() <= tol_g * Gref.abs().max().item() + 1e-6, k


def test_sigma_stack_equals_single_exact():
    """With the exact kernel the batched forward gives the single-matrix
    sigma of every matrix (same iteration, different summation order)."""
    banner("sigma_max stack vs single")
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max, sigma_max_stack
    Ws = torch.stack([planted(200, 264, seed=60 + k) for k in range(4)]).float().cuda()
    p = RunParams(kernel="exact", step=60)
    sb = sigma_max_stack(Ws, p)
    for k in range(4):
        ss = sigma_max(Ws[k].contiguous(), p)
        assert abs(sb[k].item() - ss.item()) <= 1e-6 * ss.item()


def test_sigma_stack_tuned():
    """tune=True tunes the batched shapes on runs first; sigma and the
    gradient keep the untuned accuracy against FP64 svdvals."""
    banner("sigma_max stack tuned")
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max_stack
    Ws = torch.stack([planted(384, 320, seed=80 + k, top=9.0 + k, second=5.0) for k in range(4)])
    W = Ws.float().cuda().requires_grad_(True)
    s = sigma_max_stack(W, RunParams(kernel="previousAMM", c=160, step=60, vr_period=5, seed=4), tune=True)
    s.sum().backward()
    Wq = W.detach().double().cpu()
    for k in range(4):
        ref = torch.linalg.svdvals(Wq[k])[0].item()
        assert abs(s[k].item() - ref) <= 5e-6 * ref, (k, s[k].item(), ref)
        Gref = _svd_grad(Wq[k])
        assert (W.grad[k].double().cpu() - Gref).abs().max().item() <= 2e-5 * Gref.abs().max().item() + 1e-6, k
