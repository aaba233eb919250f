"""Python host mirror of the reference's power-method interface.

The reference drives everything from `AMM.cu` through
`powerMethod_by_{exact,previous,proposed,proposed_v2}(host_memory&,
device_memory&, bool is_output)` (powerMethod_by_exact.cuh:4-7,
powerMethod_by_previous.cuh:100-103, powerMethod_by_proposed.cuh:105-108,
powerMethod_by_proposed_v2.cuh:199-202) and `square_of_A_by_X(device_memory&,
bool sanity_check)` for one sampled product.  Here `AmmContext` plays the
part of `device_memory` (it owns every work buffer; the matrix stays the
caller's) and `RunParams` that of `host_memory`'s run fields.  All compute
goes through libamm.so; torch is used only to hold device memory.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import AmmError, check  # noqa: F401

SAMPLE_SEG = 8192


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class RunParams:
    """Run fields of host_memory (host_memory.h:26-50)."""

    kernel: str = "exact"
    c: int = 0
    step: int = 10
    vr_period: int = 0
    vr_adaptive_threshold: float = 0.0
    seed: int = 0
    trial: int = 0
    sanity: bool = False
    use_graph: bool = True
    keep_stats: bool = False
    # the reference's variance_reduction = false (host_memory.h:108-120): no
    # variance-reduction exact steps; proposedAMM_v2 keeps its transition
    no_vr: bool = False

    def to_c(self) -> _lib.amm_params:
        if self.kernel not in _lib.KERNEL_NAMES:
            raise ValueError(f"'-kernel {self.kernel}' is not valid.")
        flags = 0
        if not self.use_graph:
            flags |= _lib.AMM_FLAG_NO_GRAPH
        if self.keep_stats:
            flags |= _lib.AMM_FLAG_KEEP_STATS
        if self.no_vr:
            flags |= _lib.AMM_FLAG_NO_VR
        return _lib.amm_params(
            kernel=_lib.KERNEL_NAMES[self.kernel], c=int(self.c), step=int(self.step),
            vr_period=int(self.vr_period), vr_adaptive_threshold=float(self.vr_adaptive_threshold),
            seed=int(self.seed) & 0xFFFFFFFFFFFFFFFF, trial=int(self.trial) & 0xFFFFFFFF,
            sanity=1 if self.sanity else 0, flags=flags, reserved=0)


@dataclass
class RunResult:
    time_ms: np.ndarray  # cumulative device time per iteration
    eig: np.ndarray      # raw Rayleigh quotient y^T x of A A^T per iteration
    err: np.ndarray      # y^T v* / ||y|| per iteration
    wall_ms: float       # HIP-event time of the whole run

    def eigenvalue(self, diagonal_offset: float = 0.0) -> np.ndarray:
        """Printed Eigenvalue column: max(0, sqrt(max(0, eig)) - delta)
        (powerMethod_by_exact.cuh:83)."""
        return np.maximum(0.0, np.sqrt(np.maximum(0.0, self.eig)) - diagonal_offset)

    def norm_error(self) -> np.ndarray:
        """Printed NormError column: 1 - err^2 (powerMethod_by_exact.cuh:84)."""
        return 1.0 - self.err * self.err


class AmmContext:
    """Context over a device-resident d x n column-major matrix.

    `A` is either a torch tensor on the GPU holding the matrix as an
    (n, ld) row-major array (row j = column j of A, i.e. column-major A),
    or a raw device pointer with explicit d, n, ld and dtype.
    """

    def __init__(self, A, d: int | None = None, n: int | None = None, ld: int | None = None,
                 dtype: str | None = None, stream: int = 0):
        self._lib = _lib.load()
        self._keep = A
        if hasattr(A, "data_ptr"):
            import torch
            if A.dim() != 2:
                raise ValueError("A must be 2-D (n, ld): one row per column of the matrix")
            if not A.is_cuda:
                raise ValueError("A must be a device tensor")
            if A.stride(1) != 1:
                raise ValueError("A rows must be contiguous")
            n_ = A.shape[0]
            ld_ = A.stride(0)
            d_ = A.shape[1] if d is None else d
            if A.dtype == torch.float32:
                dt = _lib.AMM_DTYPE_F32
            elif A.dtype == torch.bfloat16:
                dt = _lib.AMM_DTYPE_BF16
            else:
                raise ValueError("A must be float32 or bfloat16")
            ptr = A.data_ptr()
            d, n, ld = d_, n_, ld_
            if stream == 0:
                stream = torch.cuda.current_stream(A.device).cuda_stream
        else:
            ptr = int(A)
            dt = {"f32": _lib.AMM_DTYPE_F32, "bf16": _lib.AMM_DTYPE_BF16}[dtype or "f32"]
        self.d, self.n, self.ld, self.dtype = int(d), int(n), int(ld), dt
        h = C.c_void_p()
        check(self._lib.amm_ctx_create(C.c_void_p(ptr), dt, self.d, self.n, self.ld,
                                       C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.amm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference eigenvector (device_memory.h:246-254) --------------------
    def set_exact_eigenvector(self, v: np.ndarray | None) -> None:
        if v is None:
            check(self._lib.amm_set_exact_eigenvector(self._h, None, 0))
            return
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(self._lib.amm_set_exact_eigenvector(self._h, C.c_void_p(_ptr(v)), v.size))

    # -- get_norms (device_memory.h:34-59) ---------------------------------
    def column_stats(self):
        alpha = np.empty(self.n, np.float32)
        nrm2 = np.empty(self.n, np.float64)
        cnrm2 = np.empty(self.n, np.float64)
        a2 = np.zeros(1, np.float64)
        check(self._lib.amm_column_stats(self._h, C.c_void_p(_ptr(alpha)), C.c_void_p(_ptr(nrm2)),
                                         C.c_void_p(_ptr(cnrm2)), C.c_void_p(_ptr(a2))))
        return alpha, nrm2, cnrm2, float(a2[0])

    # -- set_weights + InclusiveSum (common.h:34-40) ------------------------
    def build_cdf(self, centred: bool = False):
        cdf = np.empty(self.n, np.uint64)
        tot = np.zeros(1, np.uint64)
        lp = np.zeros(1, np.uint64)
        check(self._lib.amm_build_cdf(self._h, 1 if centred else 0, C.c_void_p(_ptr(cdf)),
                                      C.c_void_p(_ptr(tot)), C.c_void_p(_ptr(lp))))
        return cdf, int(tot[0]), int(lp[0])

    # -- pick_index + SortKeys + update_weights (device_memory.h:519-545) ---
    def sample(self, c: int, iteration: int, seed: int, trial: int = 0, centred: bool = False):
        """Returns per-segment (idx, wgt) lists: unique sorted columns and
        their weights mult * total / (c * q_j)."""
        idx = np.zeros(c, np.uint32)
        wgt = np.zeros(c, np.float32)
        nseg = (c + SAMPLE_SEG - 1) // SAMPLE_SEG
        segc = np.zeros(nseg, np.uint32)
        check(self._lib.amm_sample(self._h, 1 if centred else 0, seed & 0xFFFFFFFFFFFFFFFF, iteration,
                                   trial, c, C.c_void_p(_ptr(idx)), C.c_void_p(_ptr(wgt)),
                                   C.c_void_p(_ptr(segc))))
        return idx, wgt, segc

    # -- square_of_A_by_previous / _proposed --------------------------------
    def square_of_A(self, kernel: str, x, c: int, iteration: int, seed: int, trial: int = 0,
                    x_memo=None, y_memo=None, sanity: bool = False, y=None):
        import torch
        k = _lib.KERNEL_NAMES[kernel]
        if y is None:
            y = torch.empty(self.d, dtype=torch.float32, device=x.device)
        for t in (x, x_memo, y_memo, y):
            if t is not None and (t.dtype != torch.float32 or t.numel() != self.d or not t.is_contiguous()):
                raise ValueError("vectors must be contiguous float32 of length d")
        p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        check(self._lib.amm_square_of_A(self._h, k, 1 if sanity else 0, c, seed & 0xFFFFFFFFFFFFFFFF,
                                        iteration, trial, p(x), p(x_memo), p(y_memo), p(y)))
        return y

    def exact_product(self, x, y=None):
        import torch
        if y is None:
            y = torch.empty(self.d, dtype=torch.float32, device=x.device)
        check(self._lib.amm_exact_product(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())))
        return y

    # -- powerMethod_by_X ----------------------------------------------------
    def power_method(self, params: RunParams) -> RunResult:
        p = params.to_c()
        S = int(params.step)
        t = np.zeros(S, np.float64)
        e = np.zeros(S, np.float64)
        r = np.zeros(S, np.float64)
        check(self._lib.amm_power_method(self._h, C.byref(p), C.c_void_p(_ptr(t)), C.c_void_p(_ptr(e)),
                                         C.c_void_p(_ptr(r))))
        ms = np.zeros(1, np.float64)
        check(self._lib.amm_last_run_ms(self._h, C.c_void_p(_ptr(ms))))
        return RunResult(t, e, r, float(ms[0]))

    def last_modes(self, step: int) -> np.ndarray:
        m = np.zeros(step, np.uint32)
        check(self._lib.amm_last_modes(self._h, C.c_void_p(_ptr(m)), step))
        return m

    def time_product(self, kernel: str, c: int = 0, seed: int = 0, iters: int = 100):
        """Mean duration (ms) of the two contraction kernels of one product
        (HIP events on the context's stream) and the columns one launch reads."""
        k = _lib.KERNEL_NAMES[kernel]
        t = np.zeros(2, np.float64)
        u = np.zeros(1, np.uint64)
        check(self._lib.amm_time_product(self._h, k, c, seed & 0xFFFFFFFFFFFFFFFF, iters,
                                         C.c_void_p(_ptr(t)), C.c_void_p(_ptr(t) + 8),
                                         C.c_void_p(_ptr(u))))
        return float(t[0]), float(t[1]), int(u[0])

    def time_column_stats(self, iters: int = 20) -> float:
        """Mean duration (ms) of the column-statistics pass (one read of A)."""
        t = np.zeros(1, np.float64)
        check(self._lib.amm_time_column_stats(self._h, iters, C.c_void_p(_ptr(t))))
        return float(t[0])

    def time_pass(self, pass_: str, iters: int = 20) -> float:
        """Mean duration (ms) of one pass as a run issues it: "stats_z"
        (column statistics with z = A alpha, kernel and partial sum) or
        "exact" (the exact product without its finish); or "read", a plain
        streaming read of the matrix, the read ceiling for the others."""
        code = {"stats_z": 1, "exact": 2, "read": 3}[pass_]
        t = np.zeros(1, np.float64)
        check(self._lib.amm_time_pass(self._h, code, iters, C.c_void_p(_ptr(t))))
        return float(t[0])

    def sigma_max(self, params: RunParams, device=None):
        """sigma_max(A) = sqrt(lambda_max(A A^T)) with unit singular vectors:
        runs the power method of `params`, u = its final iterate (length d),
        then one exact pass t = A^T u gives sigma = ||t||, v = t / sigma
        (length n).  The reference prints sqrt of the Rayleigh quotient
        (powerMethod_by_exact.cuh:83); this form also returns u and v so that
        d sigma / d A = u v^T can be applied."""
        import torch
        if device is None:
            device = self._keep.device if hasattr(self._keep, "device") else torch.device("cuda")
        p = params.to_c()
        u = torch.empty(self.d, dtype=torch.float32, device=device)
        v = torch.empty(self.n, dtype=torch.float32, device=device)
        s = np.zeros(1, np.float64)
        check(self._lib.amm_sigma_max(self._h, C.byref(p), C.c_void_p(_ptr(s)),
                                      C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr())))
        return float(s[0]), u, v

    def get_vector(self, which: str = "x") -> np.ndarray:
        v = np.empty(self.d, np.float32)
        check(self._lib.amm_get_vector(self._h, 0 if which == "x" else 1, C.c_void_p(_ptr(v)), self.d))
        return v


def rank1_update(out, g: float, u, v, d: int, n: int, ld: int) -> None:
    """out[:, j] = g * u * v[j] for the column-major d x n matrix held by the
    device tensor `out` (leading dimension ld; float32 or bfloat16): the
    backward of sigma_max (d sigma / d A = u v^T)."""
    import torch
    lib = _lib.load()
    if out.dtype == torch.float32:
        dt = _lib.AMM_DTYPE_F32
    elif out.dtype == torch.bfloat16:
        dt = _lib.AMM_DTYPE_BF16
    else:
        raise ValueError("out must be float32 or bfloat16")
    for t, m in ((u, d), (v, n)):
        if t.dtype != torch.float32 or t.numel() != m or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("u, v must be contiguous float32 device vectors of lengths d, n")
    if out.numel() < (n - 1) * ld + d:
        raise ValueError("out is smaller than a d x n matrix of leading dimension ld")
    stream = torch.cuda.current_stream(out.device).cuda_stream
    check(lib.amm_rank1_update(C.c_void_p(out.data_ptr()), dt, d, n, ld, float(g),
                               C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                               C.c_void_p(stream) if stream else None))


def rank1_update_batched(out, g, U, V, d: int, n: int, ld: int) -> None:
    """out[k][:, j] = g[k] * U[k] * V[k, j] for a stack of column-major
    d x n matrices held by the contiguous device tensor `out` of shape
    (count, n, ld) (float32 or bfloat16), in one launch; g (count,), U
    (count, d) and V (count, n) are float32 device tensors."""
    import torch
    lib = _lib.load()
    if out.dtype == torch.float32:
        dt = _lib.AMM_DTYPE_F32
    elif out.dtype == torch.bfloat16:
        dt = _lib.AMM_DTYPE_BF16
    else:
        raise ValueError("out must be float32 or bfloat16")
    count = out.shape[0]
    if out.dim() != 3 or tuple(out.shape[1:]) != (n, ld) or not out.is_contiguous():
        raise ValueError("out must be a contiguous (count, n, ld) device tensor")
    for t, shape in ((g, (count,)), (U, (count, d)), (V, (count, n))):
        if t.dtype != torch.float32 or tuple(t.shape) != shape or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("g, U, V must be contiguous float32 device tensors of shapes "
                             "(count,), (count, d), (count, n)")
    stream = torch.cuda.current_stream(out.device).cuda_stream
    check(lib.amm_rank1_update_batched(C.c_void_p(out.data_ptr()), dt, d, n, ld, count, n * ld,
                                       C.c_void_p(g.data_ptr()), C.c_void_p(U.data_ptr()), d,
                                       C.c_void_p(V.data_ptr()), n, C.c_void_p(stream) if stream else None))


def exact_product_noalloc(A, x, y=None, d: int | None = None):
    """y = A (A^T x) for the matrix held by the device tensor A ((n, ld)
    rows = columns, float32 or bfloat16) with no device allocation in the
    library (amm_exact_product_noalloc: its intermediates live in a fixed
    16 MiB device array per GPU).  Asynchronous on torch's current stream;
    y (float32, length d) is allocated here by torch when not given."""
    import torch
    lib = _lib.load()
    if A.dim() != 2 or not A.is_cuda or A.stride(1) != 1:
        raise ValueError("A must be a 2-D device tensor (n, ld) with contiguous rows")
    if A.dtype == torch.float32:
        dt = _lib.AMM_DTYPE_F32
    elif A.dtype == torch.bfloat16:
        dt = _lib.AMM_DTYPE_BF16
    else:
        raise ValueError("A must be float32 or bfloat16")
    n, ld = A.shape[0], A.stride(0)
    d = A.shape[1] if d is None else int(d)
    if y is None:
        y = torch.empty(d, dtype=torch.float32, device=A.device)
    for t in (x, y):
        if t.dtype != torch.float32 or t.numel() != d or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("x and y must be contiguous float32 device vectors of length d")
    stream = torch.cuda.current_stream(A.device).cuda_stream
    check(lib.amm_exact_product_noalloc(C.c_void_p(A.data_ptr()), dt, d, n, ld, C.c_void_p(x.data_ptr()),
                                        C.c_void_p(y.data_ptr()), C.c_void_p(stream) if stream else None))
    return y


def _run(kernel: str, ctx: AmmContext, **kw) -> RunResult:
    return ctx.power_method(RunParams(kernel=kernel, **kw))


def powerMethod_by_exact(ctx: AmmContext, **kw) -> RunResult:
    """powerMethod_by_exact.cuh:4-100"""
    return _run("exact", ctx, **kw)


def powerMethod_by_previous(ctx: AmmContext, **kw) -> RunResult:
    """powerMethod_by_previous.cuh:100-210"""
    return _run("previousAMM", ctx, **kw)


def powerMethod_by_proposed(ctx: AmmContext, **kw) -> RunResult:
    """powerMethod_by_proposed.cuh:105-235"""
    return _run("proposedAMM", ctx, **kw)


def powerMethod_by_proposed_v2(ctx: AmmContext, **kw) -> RunResult:
    """powerMethod_by_proposed_v2.cuh:199-377"""
    return _run("proposedAMM_v2", ctx, **kw)


@dataclass
class BatchResult:
    time_ms: np.ndarray  # [step] device time from the run's first kernel (all problems together)
    eig: np.ndarray      # [problems][trials][step] raw Rayleigh quotients
    err: np.ndarray      # [problems][trials][step] y^T v* / ||y||
    wall_ms: float

    def eigenvalue(self) -> np.ndarray:
        return np.sqrt(np.maximum(0.0, self.eig))


class AmmBatch:
    """Batched power method over many device matrices, `trials` (<= 16)
    sampling seeds each, stacked as one block of right-hand sides.

    `problems` is a list of dicts with keys A (a device tensor (n, ld) whose
    row j is column j of A, float32 or bfloat16), c (sampled columns per
    product, at most min(n, 65535): the engine's draw counts are 16-bit),
    seed, and optionally d (rows, default ld).  Trial k of problem
    i reproduces the single-matrix run with seed = problem seed and trial =
    params.trial + k.
    """

    def __init__(self, problems, trials: int = 16, stream: int = 0):
        import torch
        self._lib = _lib.load()
        self._keep = [p["A"] for p in problems]
        arr = (_lib.amm_problem * len(problems))()
        dev = None
        for i, p in enumerate(problems):
            A = p["A"]
            if A.dim() != 2 or not A.is_cuda or A.stride(1) != 1:
                raise ValueError("A must be a 2-D device tensor with contiguous rows")
            if A.dtype == torch.float32:
                dt = _lib.AMM_DTYPE_F32
            elif A.dtype == torch.bfloat16:
                dt = _lib.AMM_DTYPE_BF16
            else:
                raise ValueError("A must be float32 or bfloat16")
            d = int(p.get("d", A.shape[1]))
            arr[i] = _lib.amm_problem(A=A.data_ptr(), dtype=dt, c=int(p.get("c", 0)), d=d,
                                      n=int(A.shape[0]), ld=int(A.stride(0)),
                                      seed=int(p.get("seed", 0)) & 0xFFFFFFFFFFFFFFFF)
            dev = A.device
        if stream == 0:
            stream = torch.cuda.current_stream(dev).cuda_stream
        self.count = len(problems)
        self.trials = int(trials)
        self.dims = [(int(arr[i].d), int(arr[i].n)) for i in range(self.count)]
        h = C.c_void_p()
        check(self._lib.amm_batch_create(arr, self.count, self.trials,
                                         C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.amm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_exact_eigenvector(self, problem: int, v) -> None:
        if v is None:
            check(self._lib.amm_batch_set_exact_eigenvector(self._h, problem, None, 0))
            return
        a = np.ascontiguousarray(v, np.float64)
        check(self._lib.amm_batch_set_exact_eigenvector(self._h, problem, C.c_void_p(_ptr(a)), a.size))

    def power_method(self, params: RunParams) -> BatchResult:
        p = params.to_c()
        S = int(params.step)
        t = np.zeros(S, np.float64)
        e = np.zeros((self.count, self.trials, S), np.float64)
        r = np.zeros((self.count, self.trials, S), np.float64)
        check(self._lib.amm_batch_power_method(self._h, C.byref(p), C.c_void_p(_ptr(t)),
                                               C.c_void_p(_ptr(e)), C.c_void_p(_ptr(r))))
        ms = np.zeros(1, np.float64)
        check(self._lib.amm_batch_last_run_ms(self._h, C.c_void_p(_ptr(ms))))
        return BatchResult(t, e, r, float(ms[0]))

    def last_modes(self, step: int) -> np.ndarray:
        m = np.zeros((self.count, self.trials, step), np.uint32)
        check(self._lib.amm_batch_last_modes(self._h, C.c_void_p(_ptr(m)), m.size))
        return m

    def get_vector(self, problem: int, trial: int, which: str = "x") -> np.ndarray:
        v = np.empty(self.dims[problem][0], np.float32)
        check(self._lib.amm_batch_get_vector(self._h, problem, trial, 0 if which == "x" else 1,
                                             C.c_void_p(_ptr(v)), v.size))
        return v

    def set_tuning(self, problem: int, dot_blocks: int, passes: int, chunks: int) -> None:
        check(self._lib.amm_batch_set_tuning(self._h, problem, dot_blocks, passes, chunks))

    def get_tuning(self, problem: int):
        a = np.zeros(3, np.uint32)
        check(self._lib.amm_batch_get_tuning(self._h, problem, C.c_void_p(_ptr(a)),
                                             C.c_void_p(_ptr(a) + 4), C.c_void_p(_ptr(a) + 8)))
        return int(a[0]), int(a[1]), int(a[2])

    def autotune(self, iters: int = 5) -> float:
        """Shapes tuned on the dense pass; returns its time (ms)."""
        ms = np.zeros(1, np.float64)
        check(self._lib.amm_batch_autotune(self._h, iters, C.c_void_p(_ptr(ms))))
        return float(ms[0])

    def autotune_run(self, params: RunParams, reps: int = 2) -> float:
        """Shapes tuned on complete runs of `params`; returns the best run's
        device time (ms)."""
        p = params.to_c()
        ms = np.zeros(1, np.float64)
        check(self._lib.amm_batch_autotune_run(self._h, C.byref(p), reps, C.c_void_p(_ptr(ms))))
        return float(ms[0])

    def time_dense(self, iters: int = 5) -> float:
        """Mean time (ms) of one dense pass of both contractions with the
        shapes in place."""
        ms = np.zeros(1, np.float64)
        check(self._lib.amm_batch_time_dense(self._h, iters, C.c_void_p(_ptr(ms))))
        return float(ms[0])

    def sigma_max(self, params: RunParams, vectors: bool = True, out_u=None, out_v=None):
        """sigma_max of every (problem, trial) ([problems][trials] float64)
        and, for trial 0 of each problem, the unit singular vectors u
        (length d) and v (length n) as lists of device tensors, or as the
        rows of out_u (problems, d) and out_v (problems, n) when those
        contiguous float32 device tensors are given (problems of one shape)."""
        import torch
        p = params.to_c()
        sig = np.zeros((self.count, self.trials), np.float64)
        U = V = None
        up = vp = None
        if vectors and (out_u is not None or out_v is not None):
            d, n = self.dims[0]
            if any(dm != (d, n) for dm in self.dims):
                raise ValueError("out_u / out_v need problems of one shape")
            for t, m in ((out_u, d), (out_v, n)):
                if (t is None or t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous()
                        or tuple(t.shape) != (self.count, m)):
                    raise ValueError(f"out_u / out_v must be contiguous float32 device tensors of "
                                     f"shapes ({self.count}, {d}) and ({self.count}, {n})")
            U, V = out_u, out_v
            up = (C.c_void_p * self.count)(*[out_u.data_ptr() + 4 * d * i for i in range(self.count)])
            vp = (C.c_void_p * self.count)(*[out_v.data_ptr() + 4 * n * i for i in range(self.count)])
        elif vectors:
            dev = self._keep[0].device
            U = [torch.empty(d, dtype=torch.float32, device=dev) for d, _ in self.dims]
            V = [torch.empty(n, dtype=torch.float32, device=dev) for _, n in self.dims]
            up = (C.c_void_p * self.count)(*[u.data_ptr() for u in U])
            vp = (C.c_void_p * self.count)(*[v.data_ptr() for v in V])
        check(self._lib.amm_batch_sigma_max(self._h, C.byref(p), C.c_void_p(_ptr(sig)), up, vp))
        return sig, U, V

    def sample(self, kernel: str, iteration: int, trial0: int, problem: int):
        """Union of the trials' samples of one iteration: (columns ascending,
        weights [columns][trials])."""
        n = self.dims[problem][1]
        cols = np.zeros(n, np.uint32)
        w = np.zeros((n, self.trials), np.float32)
        cnt = np.zeros(1, np.uint64)
        check(self._lib.amm_batch_sample(self._h, _lib.KERNEL_NAMES[kernel], iteration, trial0, problem,
                                         C.c_void_p(_ptr(cols)), C.c_void_p(_ptr(w)),
                                         C.c_void_p(_ptr(cnt))))
        m = int(cnt[0])
        return cols[:m].copy(), w[:m].copy()
