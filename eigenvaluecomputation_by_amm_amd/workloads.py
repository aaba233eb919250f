"""Synthetic workloads (generated on the GPU with torch) and the FP64 Lanczos
ground truth used by the benchmark.

Matrices are returned in the library's layout: a (n, ld) tensor whose row j
is column j of the d x n matrix A.
"""
from __future__ import annotations

import math

import torch


def _haar(rows: int, cols: int, gen: torch.Generator, device) -> torch.Tensor:
    G = torch.randn((rows, cols), generator=gen, device=device, dtype=torch.float64)
    Q, R = torch.linalg.qr(G)
    return Q * torch.sign(torch.diagonal(R))[None, :]


def heavy_tailed(d: int = 16384, n: int = 65536, seed: int = 2024, sigma: float = 1.5,
                 rank: int = 8, strength: float = 50.0, device="cuda") -> torch.Tensor:
    """A_ij = N(0,1) * l_j with l_j ~ lognormal(0, sigma), plus a planted
    rank-`rank` term sum_k theta_k u_k v_k^T, theta_k = strength *
    (sqrt(d) + sqrt(n)) * 0.9^(k-1), u, v Haar.  FP32, (n, d) layout.
    (16384 x 65536: the single-large-matrix workload, 4 GiB.)"""
    gen = torch.Generator(device=device).manual_seed(seed)
    ell = torch.exp(sigma * torch.randn(n, generator=gen, device=device, dtype=torch.float64))
    A = torch.randn((n, d), generator=gen, device=device, dtype=torch.float32)
    A.mul_(ell.float()[:, None])
    U = _haar(d, rank, gen, device)
    V = _haar(n, rank, gen, device)
    th = strength * (math.sqrt(d) + math.sqrt(n)) * (0.9 ** torch.arange(rank, device=device,
                                                                         dtype=torch.float64))
    A.addmm_((V * th).float(), U.T.float().contiguous())
    return A


def lowrank_plus_noise(d: int = 1024, n: int = 2048, seed: int = 0, noise: float = 0.01,
                       device="cuda") -> torch.Tensor:
    """A = U diag(s) V^T + noise*G, s_i = 10/i, Haar U, V (the 1024 x 2048
    CPU-runnable workload).  FP32, (n, d) layout."""
    gen = torch.Generator(device=device).manual_seed(seed)
    r = min(d, n)
    U = _haar(d, r, gen, device)
    V = _haar(n, r, gen, device)
    s = 10.0 / torch.arange(1, r + 1, device=device, dtype=torch.float64)
    A = (V * s) @ U.T + noise * torch.randn((n, d), generator=gen, device=device, dtype=torch.float64)
    return A.float().contiguous()


@torch.no_grad()
def lambda_max_lanczos(A_cols: torch.Tensor, d: int | None = None, k: int = 64, seed: int = 0):
    """Largest eigenvalue of A A^T in FP64 by Lanczos with full
    re-orthogonalisation; A_cols is the (n, ld) layout (any float dtype).
    Returns (lambda, change over the last 8 Lanczos steps)."""
    d = A_cols.shape[1] if d is None else d
    A64 = A_cols[:, :d].double()
    dev = A64.device
    gen = torch.Generator(device=dev).manual_seed(seed)
    q = torch.randn(d, generator=gen, device=dev, dtype=torch.float64)
    q /= torch.linalg.norm(q)
    Q = torch.zeros((k + 1, d), device=dev, dtype=torch.float64)
    Q[0] = q
    al, be, hist = [], [], []
    for j in range(k):
        w = A64.T @ (A64 @ Q[j])
        a = float(Q[j] @ w)
        w -= Q[: j + 1].T @ (Q[: j + 1] @ w)
        w -= Q[: j + 1].T @ (Q[: j + 1] @ w)
        al.append(a)
        b = float(torch.linalg.norm(w))
        T = torch.diag(torch.tensor(al, dtype=torch.float64))
        if be:
            off = torch.tensor(be, dtype=torch.float64)
            T += torch.diag(off, 1) + torch.diag(off, -1)
        hist.append(float(torch.linalg.eigvalsh(T)[-1]))
        if b < 1e-300 or j == k - 1:
            break
        be.append(b)
        Q[j + 1] = w / b
    lam = hist[-1]
    change = abs(hist[-1] - hist[-9]) / lam if len(hist) > 8 else float("inf")
    del A64
    return lam, change


def uneven_problem_shapes(count: int = 256, seed: int = 7):
    """(d, n) of the many-matrix workload: d ~ U[512, 8192], n ~ U[1024,
    32768], rounded up to multiples of 8 (16-byte BF16 columns)."""
    g = torch.Generator().manual_seed(seed)
    d = torch.randint(512, 8193, (count,), generator=g)
    n = torch.randint(1024, 32769, (count,), generator=g)
    return [(int((a + 7) // 8 * 8), int((b + 7) // 8 * 8)) for a, b in zip(d.tolist(), n.tolist())]


def padded_ld(d: int, dtype=torch.bfloat16) -> int:
    """Leading dimension with every column starting on a 512-byte boundary
    (256 BF16 / 128 FP32 elements): the batched contractions stage 512-byte
    runs of each column, and a run that straddles an extra cache line costs
    about 10% of their bandwidth."""
    per = 512 // torch.tensor([], dtype=dtype).element_size()
    return (d + per - 1) // per * per


def with_ld(M: torch.Tensor, ld: int) -> torch.Tensor:
    """The (n, d) matrix M as the first d entries of the rows of an (n, ld)
    buffer (zeros beyond d): the same column-major matrix with leading
    dimension ld."""
    n, d = M.shape
    if ld == d:
        return M.contiguous()
    out = torch.zeros((n, ld), dtype=M.dtype, device=M.device)
    out[:, :d] = M
    return out


def lowrank_plus_noise_fast(d: int, n: int, seed: int, rank: int = 64, noise: float = 0.01,
                            device="cuda") -> torch.Tensor:
    """The 1024 x 2048 construction (s_i = 10/i, Haar factors, 0.01 N(0,1)
    noise) truncated to its top `rank` singular pairs: s_65 = 0.15 lies far
    below the noise edge 0.01 (sqrt d + sqrt n) >= 0.55, so the top of the
    spectrum is that of the full construction while generation costs
    O((d + n) rank^2 + d n rank) instead of a full QR.  FP32, (n, d) layout."""
    gen = torch.Generator(device=device).manual_seed(seed)
    r = min(rank, d, n)
    U = _haar(d, r, gen, device).float()
    V = _haar(n, r, gen, device).float()
    s = (10.0 / torch.arange(1, r + 1, device=device, dtype=torch.float32))
    A = torch.randn((n, d), generator=gen, device=device, dtype=torch.float32).mul_(noise)
    A.addmm_(V * s, U.T.contiguous())
    return A


@torch.no_grad()
def grf_images(side: int = 64, n: int = 40000, length_scale: float = 4.0, noise: float = 0.05,
               seed: int = 0, device="cuda", chunk: int = 4096) -> torch.Tensor:
    """Image-like columns: n images of side x side pixels (d = side^2 rows).
    Each image is a Gaussian random field with squared-exponential
    covariance exp(-r^2 / (2 l^2)) (spectral synthesis: white noise filtered
    by exp(-pi^2 l^2 |k|^2), the square root of that kernel's spectrum, on
    the periodic grid), min-max scaled to [0, 1], mixed with `noise` of
    U[0, 1] noise (x = (1 - noise) f + noise u), and finally every row
    (pixel) is centred over the images.  FP32, (n, d) layout."""
    gen = torch.Generator(device=device).manual_seed(seed)
    ky = torch.fft.fftfreq(side, device=device, dtype=torch.float64)
    kx = torch.fft.rfftfreq(side, device=device, dtype=torch.float64)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    amp = torch.exp(-(math.pi ** 2) * length_scale ** 2 * k2).float()
    out = torch.empty((n, side * side), device=device, dtype=torch.float32)
    for b in range(0, n, chunk):
        m = min(chunk, n - b)
        w = torch.randn((m, side, side), generator=gen, device=device, dtype=torch.float32)
        f = torch.fft.irfft2(torch.fft.rfft2(w) * amp, s=(side, side))
        lo = f.amin(dim=(1, 2), keepdim=True)
        hi = f.amax(dim=(1, 2), keepdim=True)
        f = (f - lo) / (hi - lo)
        u = torch.rand((m, side, side), generator=gen, device=device, dtype=torch.float32)
        out[b:b + m] = ((1.0 - noise) * f + noise * u).reshape(m, -1)
    out -= out.mean(dim=0, keepdim=True)
    return out
