"""Shared test helpers: small seeded matrices in the library's layout."""
from __future__ import annotations

import numpy as np


def haar(rng: np.random.Generator, rows: int, cols: int) -> np.ndarray:
    """rows x cols with orthonormal columns (QR of a Gaussian, sign-fixed)."""
    G = rng.standard_normal((rows, cols))
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))


def lowrank_plus_noise(d: int, n: int, seed: int = 0, noise: float = 0.01, decay: float = 1.0) -> np.ndarray:
    """A = U diag(s) V^T + noise*G, s_i = 10 i^-decay (the small CPU-runnable
    workload: 1024 x 2048 at full size).  Returns d x n float32."""
    rng = np.random.default_rng(seed)
    r = min(d, n)
    U = haar(rng, d, r)
    V = haar(rng, n, r)
    s = 10.0 * np.arange(1, r + 1, dtype=np.float64) ** (-decay)
    A = (U * s) @ V.T + noise * rng.standard_normal((d, n))
    return A.astype(np.float32)


def heavy_columns(d: int, n: int, seed: int = 0, sigma: float = 1.5, rank: int = 8,
                  strength: float = 50.0) -> np.ndarray:
    """Gaussian entries, columns scaled by lognormal(0, sigma), plus a planted
    rank-`rank` term sum_k theta_k u_k v_k^T with theta_k = strength *
    (sqrt(d) + sqrt(n)) * 0.9^(k-1).  d x n float32."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((d, n)) * rng.lognormal(0.0, sigma, size=n)[None, :]
    U = haar(rng, d, rank)
    V = haar(rng, n, rank)
    th = strength * (np.sqrt(d) + np.sqrt(n)) * 0.9 ** np.arange(rank)
    A += (U * th) @ V.T
    return A.astype(np.float32)


def to_cols(A: np.ndarray, ld: int | None = None) -> np.ndarray:
    """d x n matrix -> (n, ld) array whose row j is column j (column-major A)."""
    d, n = A.shape
    ld = d if ld is None else ld
    out = np.zeros((n, ld), dtype=A.dtype)
    out[:, :d] = A.T
    return out


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def rel(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
