"""The differentiation configuration at full size: sigma_max of a stack of
64 weight-like matrices W_k = G / sqrt(4096) + sum_{i<4} s_i u_i v_i^T
(s = 5, 4, 3, 2; Haar u, v), 4096 x 4096, BF16.  Forward: one batched AMM
power method over the stack (previousAMM, c = 1024, exact step every 5, 50
iterations) and one exact pass for v; backward: the rank-1 adjoint
g_k v_k u_k^T of every matrix in one launch.

* sigma of 4 matrices within 1e-6 of an FP64 power iteration on the stored
  values, and their gradients within the BF16 rounding of the stored
  gradient (2^-8 of its largest entry) of the textbook derivative
  g_k u_1 v_1^T;
* matrix 0: the directional derivative along 32 random unit directions
  and along the gradient itself against FP64 central differences
  (h = 1e-3) of sigma, each sigma(W +- hD) from an FP64 power iteration.
The reference prints sqrt(lambda) (powerMethod_by_exact.cuh:83) and has no
derivative; the forward loop is powerMethod_by_previous.cuh:100-210.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

B, M = 64, 4096


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def sigma64(W, x, iters):
    """FP64 power iteration on W^T W from x: sigma, right and left vectors."""
    for _ in range(iters):
        x = W.T @ (W @ x)
        x = x / torch.linalg.norm(x)
    t = W @ x
    s = torch.linalg.norm(t)
    return float(s), x, t / s


@pytest.fixture(scope="module")
def c5():
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max_stack
    from eigenvaluecomputation_by_amm_amd.workloads import _haar
    banner("C5 generate 64 x 4096x4096 BF16, sigma_max_stack forward + backward")
    gen = torch.Generator(device="cuda").manual_seed(17)
    W = torch.empty(B, M, M, dtype=torch.bfloat16, device="cuda")
    sv = torch.tensor([5.0, 4.0, 3.0, 2.0], device="cuda", dtype=torch.float64)
    for k in range(B):
        G = torch.randn(M, M, generator=gen, device="cuda", dtype=torch.float32) / math.sqrt(M)
        U = _haar(M, 4, gen, "cuda")
        V = _haar(M, 4, gen, "cuda")
        W[k] = (G + ((U * sv) @ V.T).float()).to(torch.bfloat16)
    wts = torch.rand(B, generator=gen, device="cuda", dtype=torch.float32) + 0.5
    params = RunParams(kernel="previousAMM", c=1024, step=50, vr_period=5, seed=1)
    Wg = W.clone().requires_grad_(True)
    s = sigma_max_stack(Wg, params)
    (s * wts).sum().backward()
    torch.cuda.synchronize()
    yield dict(W=W, s=s.detach().double().cpu().numpy(), grad=Wg.grad, wts=wts.double().cpu().numpy())
    del W, Wg
    torch.cuda.empty_cache()


@pytest.mark.timeout(600)
def test_c5_sigma_and_gradient_against_fp64(c5):
    banner("C5 sigma + gradient vs FP64, 4 matrices")
    gen = torch.Generator(device="cuda").manual_seed(5)
    for k in (0, 21, 42, 63):
        W64 = c5["W"][k].double()
        x0 = torch.randn(M, generator=gen, device="cuda", dtype=torch.float64)
        s, right, left = sigma64(W64, x0 / torch.linalg.norm(x0), 300)
        assert abs(c5["s"][k] - s) / s < 1e-6, (k, c5["s"][k], s)
        Gref = c5["wts"][k] * torch.outer(left, right)
        G = c5["grad"][k].double()
        assert G.shape == (M, M) and c5["grad"].dtype == torch.bfloat16
        # the stored gradient is g u v^T rounded once to BF16: relative error at
        # most 2^-8 per entry (half an ulp of 7 stored mantissa bits), plus
        # the FP32 singular vectors' own error
        err = float((G - Gref).abs().max() / Gref.abs().max())
        assert err < 2.0**-8 + 2e-4, (k, err)


@pytest.mark.timeout(600)
def test_c5_gradient_finite_differences(c5):
    banner("C5 32-direction central differences, matrix 0")
    W64 = c5["W"][0].double()
    gen = torch.Generator(device="cuda").manual_seed(6)
    x0 = torch.randn(M, generator=gen, device="cuda", dtype=torch.float64)
    s0, right, left = sigma64(W64, x0 / torch.linalg.norm(x0), 300)
    G = c5["grad"][0].double() / c5["wts"][0]  # d sigma_0 / d W_0 as the library wrote it (BF16)
    Gref = torch.outer(left, right)
    h = 1e-3
    dirs = [torch.randn(M, M, generator=gen, device="cuda", dtype=torch.float64) for _ in range(32)]
    dirs.append(G.clone())
    scale = []
    for i, D in enumerate(dirs):
        D = D / torch.linalg.norm(D)
        sp, _, _ = sigma64(W64 + h * D, right, 40)
        sm, _, _ = sigma64(W64 - h * D, right, 40)
        fd = (sp - sm) / (2 * h)
        an = float((Gref * D).sum())
        lib = float((G * D).sum())
        # O(h^2) truncation of the central difference: about h^2 / 6 / gap
        assert abs(fd - an) <= 5e-7 + 1e-3 * abs(fd), (i, fd, an)
        # the library's BF16 gradient adds its rounding, about 2^-9 / sqrt(M^2)
        assert abs(lib - fd) <= 2e-6 + 1e-2 * abs(fd), (i, fd, lib)
        scale.append(abs(fd))
    # along the gradient itself the derivative is ||G|| = 1
    assert abs(scale[-1] - 1.0) < 1e-3
    assert np.median(scale[:-1]) > 1e-5  # the random directions are not degenerate
