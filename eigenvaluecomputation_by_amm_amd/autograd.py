"""Differentiable sigma_max(W) on top of the AMM power method.

sigma_max(W) = sqrt(lambda_max(W W^T)) is what the reference prints as its
Eigenvalue column (powerMethod_by_exact.cuh:83, the square root of the
Rayleigh quotient).  `sigma_max` makes it a torch autograd op: the forward
pass runs the power method in libamm.so (exact or AMM, any variance-reduction
schedule of `RunParams`) and one exact pass for v = W^T u / sigma; the
backward pass is the rank-1 adjoint d sigma / d W = u v^T written by
amm_rank1_update, with no dense SVD anywhere.

Layout: a row-major torch tensor W of shape (rows, cols) is, byte for byte,
the column-major matrix A = W^T of size cols x rows with leading dimension
cols.  sigma_max(A) = sigma_max(W); the power method on A A^T = W^T W yields
the right singular vector of W (length cols) as u_A, and v_A = A^T u_A /
sigma (length rows) is W's left singular vector.  dW = g * v_A u_A^T as a
row-major (rows, cols) array is exactly dA = g * u_A v_A^T column-major.
"""
from __future__ import annotations

import weakref
from collections import OrderedDict

import torch

from .api import AmmBatch, AmmContext, RunParams, rank1_update, rank1_update_batched

# Contexts and batches are cached per caller tensor (pointer, shape, layout,
# dtype, launch parameters, stream): they hold only the matrix's address and
# every run re-reads the matrix.  The cache keeps a weak reference to the
# caller's tensor and no strong one: when that tensor is collected its
# entry is closed, so a cached entry never keeps a weight stack alive.  At
# most _CACHE_MAX entries of each kind are kept (least recently used first
# out).  A matrix whose layout needs a padded copy is not cached: the copy
# lives for one forward call, with a context (or batch) of its own.
_CACHE_MAX = 4
_contexts: "OrderedDict[tuple, tuple]" = OrderedDict()


def _evict(cache, key, ref) -> None:
    ent = cache.get(key)
    if ent is not None and ent[0] is ref:
        del cache[key]
        ent[1].close()


def _cached(cache, key, owner: torch.Tensor, make):
    """The cached object for `owner` under `key`, made by `make()` on a miss."""
    ent = cache.get(key)
    if ent is not None and ent[0]() is not None:
        cache.move_to_end(key)
        return ent[1]
    if ent is not None:
        del cache[key]
        ent[1].close()
    obj = make()
    obj._keep = None  # the caller's tensor keeps the matrix alive, not the cache
    ref = weakref.ref(owner)
    cache[key] = (ref, obj)
    weakref.finalize(owner, _evict, cache, key, ref)
    while len(cache) > _CACHE_MAX:
        _, (_, old) = cache.popitem(last=False)
        old.close()
    return obj


def _context(owner: torch.Tensor, Wc: torch.Tensor, cols: int) -> AmmContext:
    stream = torch.cuda.current_stream(Wc.device).cuda_stream
    key = (Wc.data_ptr(), tuple(Wc.shape), Wc.stride(0), cols, Wc.dtype, Wc.device.index, stream)
    return _cached(_contexts, key, owner, lambda: AmmContext(Wc, d=cols))


def _as_columns(W: torch.Tensor) -> tuple[torch.Tensor, int]:
    """W (rows, cols) as the (rows, ld) array of a column-major cols x rows
    matrix: rows contiguous, ld a multiple of 16 bytes (a padded device copy
    when W's own layout does not qualify)."""
    rows, cols = W.shape
    width = 4 if W.dtype == torch.float32 else 8
    ok = (W.stride(1) == 1 and W.stride(0) >= cols and W.stride(0) % width == 0
          and W.data_ptr() % 16 == 0)
    if ok:
        return W, cols
    ld = (cols + width - 1) // width * width
    Wp = torch.zeros(rows, ld, dtype=W.dtype, device=W.device)
    Wp[:, :cols] = W
    return Wp, cols


class SigmaMax(torch.autograd.Function):
    """sigma = sigma_max(W) with d sigma / d W = v u^T (row-major view)."""

    @staticmethod
    def forward(ctx, W: torch.Tensor, params: RunParams):
        if not W.is_cuda or W.dim() != 2:
            raise ValueError("W must be a 2-D device tensor")
        if W.dtype not in (torch.float32, torch.bfloat16):
            raise ValueError("W must be float32 or bfloat16")
        Wd = W.detach()
        Wc, cols = _as_columns(Wd)
        if Wc.data_ptr() == Wd.data_ptr():
            sigma, u, v = _context(W, Wc, cols).sigma_max(params, device=W.device)
        else:  # padded copy: a context for this call only
            with AmmContext(Wc, d=cols) as amm:
                sigma, u, v = amm.sigma_max(params, device=W.device)
        ctx.save_for_backward(u, v)
        ctx.shape = tuple(W.shape)
        ctx.dtype = W.dtype
        return torch.tensor(sigma, dtype=torch.float32, device=W.device)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        u, v = ctx.saved_tensors
        rows, cols = ctx.shape
        dW = torch.empty(rows, cols, dtype=ctx.dtype, device=u.device)
        rank1_update(dW, float(g.item()), u, v, d=cols, n=rows, ld=cols)
        return dW, None


def sigma_max(W: torch.Tensor, params: RunParams | None = None) -> torch.Tensor:
    """Largest singular value of W (rows, cols) by the power method on
    W^T W; differentiable in W.  Default: 200 exact iterations."""
    if params is None:
        params = RunParams(kernel="exact", step=200)
    return SigmaMax.apply(W, params)


_batches: "OrderedDict[tuple, tuple]" = OrderedDict()


def _make_batch(mats, cols: int, params: RunParams, tune: bool) -> AmmBatch:
    probs = [dict(A=m, d=cols, c=int(params.c), seed=int(params.seed) + i) for i, m in enumerate(mats)]
    b = AmmBatch(probs, trials=1)
    if tune:
        b.autotune_run(params, reps=1)
    return b


def _batch(owner: torch.Tensor, mats, cols: int, params: RunParams, tune: bool = False) -> AmmBatch:
    """Batched contexts cached per (caller stack, pointers, shapes, dtype,
    c, seed, stream); with `tune`, a new one first tunes its launch shapes
    on runs of `params` (once, when it is created)."""
    stream = torch.cuda.current_stream(mats[0].device).cuda_stream
    key = (tuple(m.data_ptr() for m in mats), tuple(mats[0].shape), mats[0].stride(0), cols,
           mats[0].dtype, int(params.c), int(params.seed), stream)
    return _cached(_batches, key, owner, lambda: _make_batch(mats, cols, params, tune))


class SigmaMaxStack(torch.autograd.Function):
    """sigma_k = sigma_max(W[k]) for a stack W (B, rows, cols), all matrices
    in one batched run (one grouped launch per kernel); d sigma_k / d W[k] =
    v_k u_k^T (row-major view), all B adjoints in one launch.  Matrix k uses
    the Philox key params.seed + k."""

    @staticmethod
    def forward(ctx, W: torch.Tensor, params: RunParams, tune: bool = False):
        if not W.is_cuda or W.dim() != 3:
            raise ValueError("W must be a 3-D device tensor (stack, rows, cols)")
        if W.dtype not in (torch.float32, torch.bfloat16):
            raise ValueError("W must be float32 or bfloat16")
        B, rows, cols = W.shape
        Wd = W.detach()
        width = 4 if W.dtype == torch.float32 else 8
        U = torch.empty(B, cols, dtype=torch.float32, device=W.device)
        V = torch.empty(B, rows, dtype=torch.float32, device=W.device)
        if (Wd.stride(2) == 1 and Wd.stride(1) >= cols and Wd.stride(1) % width == 0
                and Wd.stride(0) % width == 0 and Wd.data_ptr() % 16 == 0):
            # every matrix already qualifies as is
            sig, _, _ = _batch(W, list(Wd.unbind(0)), cols, params, tune).sigma_max(params, out_u=U, out_v=V)
        else:  # padded copies: a batch for this call only
            mats = [_as_columns(Wd[k])[0] for k in range(B)]
            with _make_batch(mats, cols, params, tune) as b:
                sig, _, _ = b.sigma_max(params, out_u=U, out_v=V)
        ctx.save_for_backward(U, V)
        ctx.shape = (B, rows, cols)
        ctx.dtype = W.dtype
        return torch.tensor(sig[:, 0], dtype=torch.float32, device=W.device)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        U, V = ctx.saved_tensors
        B, rows, cols = ctx.shape
        dW = torch.empty(B, rows, cols, dtype=ctx.dtype, device=U.device)
        gd = g.detach().to(device=U.device, dtype=torch.float32).contiguous()
        rank1_update_batched(dW, gd, U.contiguous(), V.contiguous(), d=cols, n=rows, ld=cols)
        return dW, None, None


def sigma_max_stack(W: torch.Tensor, params: RunParams | None = None, tune: bool = False) -> torch.Tensor:
    """Largest singular value of every matrix of W (B, rows, cols),
    differentiable in W.  Default: 200 exact iterations.  tune: on the first
    call for this stack, tune the batched launch shapes on runs of `params`
    (about 30 runs, once; later calls reuse the cached batch)."""
    if params is None:
        params = RunParams(kernel="exact", step=200)
    return SigmaMaxStack.apply(W, params, tune)
