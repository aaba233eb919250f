#!/usr/bin/env python3
"""Time to lambda_max within 1e-3 relative error: AMM vs the exact power
method on one MI355X.

Headline workload: A in R^{16384 x 65536} FP32 (4 GiB), i.i.d. N(0,1)
entries with each column scaled by a lognormal(0, 1.5) factor, plus a
planted rank-8 term (theta_k = 50 (sqrt d + sqrt n) 0.9^(k-1)).  Ground
truth lambda* of A A^T: FP64 Lanczos on the GPU.

A "step" is one complete power-method run (200 iterations, one sampling
seed) of the headline AMM configuration (previousAMM, c = 1024, exact step
every 4 iterations); its time to tolerance is the cumulative device time
(device wall clock, from the run's first kernel, including column
statistics, CDF and sampling) at the first iteration after which the
Rayleigh quotient stays within 1e-3 of lambda*.  `value` is the median over
the timed steps.  The FP32 exact power method on the same GPU is timed the
same way and reported beside it.

The same JSON line carries
* config.c2_table: previousAMM / proposedAMM / proposedAMM_v2 at
  c = 1024, 4096, 16384 (exact step every 4), `--table-seeds` seeds each;
* roofline: the dominant kernel of the headline run against 8 TB/s;
* cpu_baseline: the oracle port and the OpenMP exact method on the host
  cores (one warm-up, then the median of 5), numpy FP64 eigen-solves;
* batched (16 seeds of the headline matrix), c3 (image-like BF16), c4 (256
  uneven matrices x 16 seeds) and c5 (sigma_max of 64 BF16 weight
  matrices): each with its time per (matrix, seed) solve and per distinct
  matrix, the FP32 exact comparator and the same engine's exact run.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sweep] [--no-configs]
    python bench.py --config {batch,bf16,sigma,seeds,probe}   (one section, its own line)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

METRIC = "time to λ_max within 1e-3 rel err; AMM vs exact power method, 1 MI355X"
TOL = 1e-3
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch on the 16384 x 65536 FP32 workload from two rocprofv3
# PMC passes (FETCH_SIZE, then WRITE_SIZE, over `python3 bench.py --config
# probe`), each scaled by what the counter reads for a 1 GiB device copy;
# used for roofline.traffic.
PMC_TRAFFIC_16384x65536 = {
    "k_gram<float,16> (exact product A A^T x, one read of A)": 4.3123e9,
    "k_colstats<float> (column statistics)": 4.2967e9,
    "k_gather_dot<float,4,dense> (exact pass A^T x)": 4.2957e9,
    "k_gather_axpy<float,dense> (exact pass A t)": 4.3012e9,
}

# rows up to which the exact product is one column-resident pass (FP32)
GRAM_MAX_ROWS = 16384


def gram_L(d: int) -> int:
    q = -(-(-(-d // 4)) // 256) * 256
    L = 1
    while L < q // 256:
        L *= 2
    return L


# headline AMM configuration: previousAMM at c = 1024, the smallest sample
# count of the stated c set {1024, 4096, 16384}, with an exact step every 4
# iterations (the fastest period of the --sweep table)
HEADLINE = dict(kernel="previousAMM", c=1024, vr_period=4)
TABLE_KERNELS = ("previousAMM", "proposedAMM", "proposedAMM_v2")
TABLE_C = (1024, 4096, 16384)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def settled_index(eig: np.ndarray, lam: float, tol: float = TOL) -> int:
    """First i with |eig_j - lam| <= tol*lam for every j >= i (-1: never)."""
    ok = np.abs(eig - lam) <= tol * lam
    if not ok[-1]:
        return -1
    bad = np.nonzero(~ok)[0]
    return 0 if bad.size == 0 else int(bad[-1] + 1)


def first_index(eig: np.ndarray, lam: float, tol: float = TOL) -> int:
    ok = np.nonzero(np.abs(eig - lam) <= tol * lam)[0]
    return int(ok[0]) if ok.size else -1


def run_config(ctx, RunParams, lam, seeds, step, **cfg):
    out = []
    for s in seeds:
        r = ctx.power_method(RunParams(step=step, seed=s, **cfg))
        i = settled_index(r.eig, lam)
        f = first_index(r.eig, lam)
        out.append(dict(seed=s, t_settled=(r.time_ms[i] if i >= 0 else None), i_settled=i,
                        t_first=(r.time_ms[f] if f >= 0 else None), i_first=f, wall_ms=r.wall_ms,
                        run_ms=float(r.time_ms[-1]), final_rel=float(abs(r.eig[-1] - lam) / lam)))
    return out


def summarize(rows):
    ts = [r["t_settled"] for r in rows if r["t_settled"] is not None]
    return dict(median_ms=(statistics.median(ts) if ts else None),
                mean_ms=(float(np.mean(ts)) if ts else None),
                p10_ms=(float(np.percentile(ts, 10)) if ts else None),
                p90_ms=(float(np.percentile(ts, 90)) if ts else None),
                not_reached=1.0 - len(ts) / len(rows),
                median_first_ms=statistics.median([r["t_first"] for r in rows if r["t_first"] is not None])
                if any(r["t_first"] is not None for r in rows) else None,
                median_iters=statistics.median([r["i_settled"] for r in rows if r["i_settled"] >= 0])
                if ts else None,
                median_run_ms=statistics.median([r["run_ms"] for r in rows]),
                final_rel_err_median=statistics.median([r["final_rel"] for r in rows]),
                final_rel_err_max=max(r["final_rel"] for r in rows))


def rnd(v, k=4):
    return None if v is None else round(float(v), k)


def ratio(a, b, k=3):
    return round(a / b, k) if (a and b) else None


def timed(fn, reps=5, budget_s=30.0):
    """One warm-up call, then up to `reps` timed calls while the timed total
    stays under budget_s: (median ms, timed calls, last result)."""
    out = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
        if sum(ts) > budget_s:
            break
    return statistics.median(ts) * 1e3, len(ts), out


# ---------------------------------------------------------------------------
# 16384 x 65536: the C2 table and the CPU baseline
# ---------------------------------------------------------------------------
def c2_table(ctx, RunParams, lam, seeds, iters):
    """previousAMM, proposedAMM and proposedAMM_v2 at c = 1024, 4096, 16384
    with an exact step every 4 iterations over `seeds`: time to settle within
    1e-3 (median, p10, p90, fraction not reached), the median settling
    iteration and the final relative error against the FP64 value."""
    cells = {}
    for kern in TABLE_KERNELS:
        for c in TABLE_C:
            rows = run_config(ctx, RunParams, lam, seeds, iters, kernel=kern, c=c, vr_period=4)
            sm = summarize(rows)
            fr = np.array([r["final_rel"] for r in rows])
            cells[f"{kern} c={c}"] = dict(
                median_ms=rnd(sm["median_ms"]), p10_ms=rnd(sm["p10_ms"]), p90_ms=rnd(sm["p90_ms"]),
                not_reached=sm["not_reached"], median_iters=sm["median_iters"],
                final_rel_err_median=float(np.median(fr)), final_rel_err_max=float(fr.max()), seeds=len(rows))
            log(f"table {kern:15s} c={c:6d}: {cells[f'{kern} c={c}']}")
    return cells


def cpu_baseline(A_cols_host, d, lam, cfg, seed, amm_iters, exact_seed, exact_iters, args):
    """The oracle (C port of the same algorithm, OpenMP over columns and row
    blocks, every sum in its serial order) on the host cores, one warm-up
    then the median of 5: the headline AMM configuration cut at the
    iteration where the GPU run of the same seed settled (same samples, same
    trajectory), and the exact method (two passes over A per iteration) cut
    likewise.  Beside them the reference's numpy path."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    cores = O.num_threads()
    step = max(1, amm_iters + 1)
    ms, k, r = timed(lambda: O.power_method(A_cols_host, d, seed=seed, step=step, **cfg), reps=5, budget_s=20)
    rel_err = abs(r["eig"][-1] - lam) / lam
    xstep = max(1, exact_iters + 1)
    xms, xk, xr = timed(lambda: O.power_method(A_cols_host, d, kernel="exact", seed=exact_seed, step=xstep),
                        reps=5, budget_s=20)
    out = dict(value=round(ms, 2), unit="ms", cores=cores, kind="port", protocol=f"1 warm-up + median of {k}",
               sample=(f"oracle C power method, OpenMP over {cores} threads ({cfg['kernel']}, c={cfg.get('c')}, "
                       f"vr={cfg.get('vr_period')}), seed {seed}, {step} iterations (where the GPU run settled) "
                       f"on the full {d}x{A_cols_host.shape[0]} matrix incl. column statistics; "
                       f"final rel err {rel_err:.2e}"),
               exact_openmp=dict(ms=round(xms, 2), protocol=f"1 warm-up + median of {xk}", iterations=xstep,
                                 final_rel_err=float(abs(xr["eig"][-1] - lam) / lam),
                                 sample=f"oracle exact power method (A (A^T x), two passes over A per "
                                        f"iteration), seed {exact_seed}, {xstep} iterations"))
    out["numpy"] = numpy_baseline(A_cols_host, args)
    return out


def numpy_baseline(A_cols_host, args):
    """The reference's numpy path on the same host: one FP32 exact power
    iteration y = A (A^T x) with numpy's BLAS on the full host matrix, the
    FP64 eigh of A A^T that calculate_largest_eigenvalue.py runs on the
    1024 x 2048 configuration, FP32 numpy power methods on it, and (unless
    --no-configs) FP64 eigvalsh of the image-like configuration's 4096 x 4096
    Gram matrix and FP64 svdvals of one 4096 x 4096 weight matrix."""
    try:
        from threadpoolctl import threadpool_info
        blas_threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:
        blas_threads = None
    rng = np.random.default_rng(0)
    x = rng.standard_normal(A_cols_host.shape[1]).astype(np.float32)

    def exact_iter():
        t = A_cols_host @ x            # A^T x (row j of A_cols_host is column j of A)
        y = A_cols_host.T @ t          # A t
        return (y / np.linalg.norm(y)).astype(np.float32)

    it_ms, it_k, _ = timed(exact_iter, reps=5, budget_s=10)
    d1, n1 = 1024, 2048
    U, _ = np.linalg.qr(rng.standard_normal((d1, d1)))
    V, _ = np.linalg.qr(rng.standard_normal((n1, d1)))
    M = ((U * (10.0 / np.arange(1, d1 + 1))) @ V.T + 0.01 * rng.standard_normal((d1, n1))).astype(np.float32)
    G1 = M.astype(np.float64) @ M.T.astype(np.float64)
    eigh_ms, eigh_k, w = timed(lambda: np.linalg.eigh(G1)[0], reps=5, budget_s=10)
    out = dict(exact_iteration_ms=round(it_ms, 2), exact_iteration_protocol=f"1 warm-up + median of {it_k}",
               eigh_1024x2048_ms=round(eigh_ms, 2), eigh_protocol=f"1 warm-up + median of {eigh_k}",
               eigh_lambda_max=float(w[-1]), power_1024x2048=numpy_power_c1(M, float(w[-1]), rng),
               blas_threads=blas_threads)
    if not args.no_configs:
        import torch
        from eigenvaluecomputation_by_amm_amd.workloads import grf_images
        A3 = grf_images(side=64, n=40000).double()
        G3 = (A3.T @ A3).cpu().numpy()
        del A3
        e3_ms, e3_k, w3 = timed(lambda: np.linalg.eigvalsh(G3), reps=5, budget_s=30)
        out["eigvalsh_4096x40000_gram_ms"] = round(e3_ms, 1)
        out["eigvalsh_4096x40000_protocol"] = f"1 warm-up + median of {e3_k} (Gram formed on the GPU)"
        gen = torch.Generator(device="cuda").manual_seed(17)
        W = (torch.randn(4096, 4096, generator=gen, device="cuda", dtype=torch.float64) / 64.0).cpu().numpy()
        s5_ms, s5_k, _ = timed(lambda: np.linalg.svd(W, compute_uv=False), reps=5, budget_s=30)
        out["svdvals_4096x4096_ms"] = round(s5_ms, 1)
        out["svdvals_4096x4096_protocol"] = f"1 warm-up + median of {s5_k}"
        torch.cuda.empty_cache()
    out["sample"] = ("numpy FP32 A (A^T x) on the full 16384x65536 matrix; numpy FP64 eigh(A A^T) of the "
                     "1024x2048 configuration; numpy FP32 power methods (exact, previous AMM c=256) on it, 100 "
                     "iterations each; FP64 eigvalsh of the 4096x4096 Gram of the image-like configuration; "
                     "FP64 svdvals of a 4096x4096 Gaussian weight matrix")
    return out


def numpy_power_c1(M, lam, rng, iters=100, c=256):
    """A numpy power method on the 1024 x 2048 configuration, exact and
    previous AMM (c draws with p_j ~ ||a_j||^2, weights 1/(c p_j)), from one
    x0 ~ N(0, I): wall time of the 100 iterations and the final relative error
    of the Rayleigh quotient against the FP64 eigh value."""
    d1, n1 = M.shape
    x0 = rng.standard_normal(d1).astype(np.float32)
    x0 /= np.linalg.norm(x0)
    p = (M.astype(np.float64) ** 2).sum(axis=0)
    p /= p.sum()
    cdf = np.cumsum(p)
    Mt = np.ascontiguousarray(M.T)  # row j = column j: contiguous gathers
    out = {}
    for name in ("exact", "previousAMM"):
        x = x0.copy()
        lam_i = 0.0
        t0 = time.perf_counter()
        for _ in range(iters):
            if name == "exact":
                y = M @ (M.T @ x)
            else:
                j = np.minimum(np.searchsorted(cdf, rng.random(c) * cdf[-1], side="right"), n1 - 1)
                S = Mt[j]
                y = S.T @ ((1.0 / (c * p[j])).astype(np.float32) * (S @ x))
            lam_i = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
            x = (y / np.linalg.norm(y)).astype(np.float32)
        dt = time.perf_counter() - t0
        out[name] = dict(ms=round(dt * 1e3, 2), final_rel_err=abs(lam_i - lam) / lam)
    return out


# ---------------------------------------------------------------------------
# Batched engine: 16 seeds of the headline matrix
# ---------------------------------------------------------------------------
def batched_trials(A, d, n, lam_f32, iters, cfg, reps, exact_ms, amm_ms, trials=16):
    """The batched engine on the same matrix: `trials` seeds per run stacked
    as one right-hand-side block (FP32 and BF16 storage).  Per run, the time
    until every trial has settled within tolerance (device clock, setup
    included): that is the time to lambda of this one matrix
    (per_matrix_ms), and its share per (matrix, seed) solve."""
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import lambda_max_lanczos
    out = {"trials": trials, "accounting": "per_solve = per (matrix, seed); per_matrix = until all "
                                          "seeds of the one matrix have settled"}
    Ab = A.to(torch.bfloat16)
    lam_b, _ = lambda_max_lanczos(Ab, d, k=64)
    for name, M, lam, s_A in (("bf16", Ab, lam_b, 2), ("f32", A, lam_f32, 4)):
        with AmmBatch([dict(A=M, c=cfg["c"], seed=1)], trials=trials) as b:
            # shapes tuned on short runs (two VR periods) of the same configuration
            b.autotune_run(RunParams(step=2 * cfg["vr_period"], trial=20_000, **cfg), reps=2)
            dense_ms = b.time_dense(3)
            b.power_method(RunParams(step=iters, trial=10_000, **cfg))  # graph build + warm-up
            alls, nr = [], 0
            for rep in range(reps):
                r = b.power_method(RunParams(step=iters, trial=trials * rep, **cfg))
                idx = [settled_index(r.eig[0, k], lam) for k in range(trials)]
                if min(idx) < 0:
                    nr += sum(1 for i in idx if i < 0)
                    continue
                alls.append(float(r.time_ms[max(idx)]))
            rx = b.power_method(RunParams(kernel="exact", step=min(iters, 40), trial=0))
            idx = [settled_index(rx.eig[0, k], lam) for k in range(trials)]
            ex_all = float(rx.time_ms[max(idx)]) if min(idx) >= 0 else None
        t_all = statistics.median(alls) if alls else None
        out[name] = dict(
            lambda_star=lam, per_matrix_ms=rnd(t_all), per_solve_ms=rnd(t_all / trials if t_all else None, 5),
            not_reached=nr / (reps * trials),
            exact16_per_matrix_ms=rnd(ex_all), exact16_per_solve_ms=rnd(ex_all / trials if ex_all else None, 5),
            dense_pass_ms=round(dense_ms, 4),
            dense_pass_GBs=round(2 * d * n * s_A / (dense_ms * 1e-3) / 1e9, 1),
            dense_pass_frac_of_8TBs=round(2 * d * n * s_A / (dense_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
            exact_f32_vs_per_solve=ratio(exact_ms, t_all / trials if t_all else None, 2),
            exact_f32_vs_per_matrix=ratio(exact_ms, t_all, 3),
            single_amm_f32_vs_per_matrix=ratio(amm_ms, t_all, 3))
        log(f"batched {name}: {out[name]}")
    del Ab
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------
# C3: image-like BF16
# ---------------------------------------------------------------------------
def c3_section(args, reps=3):
    """40000 Gaussian-random-field images of 64 x 64 pixels (d = 4096),
    stored in BF16: proposedAMM c = 2048 (exact step every 5) on the batched
    engine with 16 seeds, and on the single-vector path (one seed), against
    the FP64 eigvalsh of the stored matrix; the FP32 exact power method on
    the FP32 data and the batched engine's own 16-seed exact run beside it."""
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import grf_images
    A32 = grf_images(side=64, n=40000)
    Ab = A32.to(torch.bfloat16)
    d, n = A32.shape[1], A32.shape[0]

    def lam64(M):
        M64 = M.double()
        G = M64.T @ M64
        del M64
        return float(torch.linalg.eigvalsh(G)[-1])

    lam_b, lam_32 = lam64(Ab), lam64(A32)
    iters = args.iterations
    cfg = dict(kernel="proposedAMM", c=2048, vr_period=5)
    with AmmContext(A32) as ctx:
        rx = [ctx.power_method(RunParams(kernel="exact", step=iters, seed=s)) for s in range(10)]
    ex_ms = statistics.median(float(r.time_ms[i]) for r in rx for i in [settled_index(r.eig, lam_32)] if i >= 0)
    with AmmContext(Ab) as ctx:
        ctx.power_method(RunParams(step=iters, seed=99, **cfg))
        single = summarize(run_config(ctx, RunParams, lam_b, list(range(10)), iters, **cfg))
    with AmmBatch([dict(A=Ab, c=cfg["c"], seed=3)], trials=16) as b:
        b.autotune_run(RunParams(step=2 * cfg["vr_period"], trial=20_000, **cfg), reps=2)
        dense_ms = b.time_dense(3)
        b.power_method(RunParams(step=iters, trial=10_000, **cfg))
        alls, fin = [], []
        for rep in range(reps):
            r = b.power_method(RunParams(step=iters, trial=16 * rep, **cfg))
            idx = [settled_index(r.eig[0, k], lam_b) for k in range(16)]
            if min(idx) >= 0:
                alls.append(float(r.time_ms[max(idx)]))
            fin.extend(float(e) for e in r.eig[0, :, -1])
        rx16 = b.power_method(RunParams(kernel="exact", step=40, trial=0))
        idx = [settled_index(rx16.eig[0, k], lam_b) for k in range(16)]
        ex16 = float(rx16.time_ms[max(idx)]) if min(idx) >= 0 else None
    fin = np.array(fin)
    t_all = statistics.median(alls) if alls else None
    gbs = 2 * d * n * 2 / (dense_ms * 1e-3) / 1e9
    out = {
        "workload": f"{n} Gaussian-random-field images 64x64 (d={d}), SE length 4 px, 5% uniform noise, "
                    "rows centred; BF16 storage",
        "kernel": cfg["kernel"], "c": cfg["c"], "vr_period": cfg["vr_period"], "iterations": iters,
        "accounting": "per_solve = per (matrix, seed); per_matrix = time to lambda of the one matrix",
        "batched16_per_matrix_ms": rnd(t_all), "batched16_per_solve_ms": rnd(t_all / 16 if t_all else None, 5),
        "single_amm_bf16_per_matrix_ms": rnd(single["median_ms"]),
        "single_amm_bf16_not_reached": single["not_reached"],
        "exact_f32_per_matrix_ms": rnd(ex_ms),
        "exact16_bf16_per_matrix_ms": rnd(ex16), "exact16_bf16_per_solve_ms": rnd(ex16 / 16 if ex16 else None, 5),
        "exact_f32_vs_batched16_per_solve": ratio(ex_ms, t_all / 16 if t_all else None, 2),
        "exact_f32_vs_batched16_per_matrix": ratio(ex_ms, t_all, 3),
        "exact_f32_vs_single_amm_per_matrix": ratio(ex_ms, single["median_ms"], 3),
        "exact16_vs_batched16_same_engine": ratio(ex16, t_all, 3),
        "final_rel_err_vs_fp64_of_bf16_max": float(np.max(np.abs(fin - lam_b) / lam_b)),
        "bf16_vs_fp32_matrix_lambda_rel": abs(lam_b - lam_32) / lam_32,
        "dense_pass_ms": round(dense_ms, 4), "dense_pass_GBs": round(gbs, 1),
        "dense_pass_frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 3),
    }
    del A32, Ab
    torch.cuda.empty_cache()
    log(f"c3: {out}")
    return out


# ---------------------------------------------------------------------------
# C4: many uneven matrices x 16 seeds
# ---------------------------------------------------------------------------
def c4_section(args):
    """256 matrices with d ~ U[512, 8192] and n ~ U[1024, 32768] in BF16,
    16 seeds each, previousAMM c = n/8 (exact step every 3), 150 iterations
    in one batched run; beside it the same with one seed per matrix, the
    batched engine's 16-seed exact run, and the FP32 exact power method run
    matrix by matrix."""
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import (lambda_max_lanczos, lowrank_plus_noise_fast, padded_ld,
                                                            uneven_problem_shapes, with_ld)
    shapes = uneven_problem_shapes(args.count)
    P = len(shapes)
    elems = sum(d * n for d, n in shapes)
    log(f"c4: generating {P} matrices, {elems / 1e9:.2f} G elements")
    m16, lam16, lam32 = [], [], []
    t_exact = []
    # columns start on 512-byte boundaries (leading dimension padded): the
    # batched contractions read 512-byte runs of each column
    for i, (d, n) in enumerate(shapes):
        M = lowrank_plus_noise_fast(d, n, seed=1000 + i)
        lam32.append(lambda_max_lanczos(M, d, k=40)[0])
        M32 = with_ld(M, padded_ld(d, torch.float32))
        with AmmContext(M32, d=d) as ctx:  # comparator: FP32 exact power method on this matrix
            rx = ctx.power_method(RunParams(kernel="exact", step=args.exact_iters, seed=5))
            j = settled_index(rx.eig, lam32[-1])
            t_exact.append(float(rx.time_ms[j]) if j >= 0 else float("nan"))
        del M32
        Mb = M.to(torch.bfloat16)
        del M
        lam16.append(lambda_max_lanczos(Mb, d, k=40)[0])
        m16.append(with_ld(Mb, padded_ld(d, torch.bfloat16)))
        del Mb
    lam16 = np.array(lam16)
    tot_exact = float(np.nansum(t_exact))
    iters = args.batch_iters
    cfg = dict(kernel="previousAMM", vr_period=args.batch_vr)
    probs = [dict(A=m16[i], d=d, c=max(1, n // 8), seed=77 + i) for i, (d, n) in enumerate(shapes)]

    def settle_all(r, ntr):
        idx = np.array([[settled_index(r.eig[i, k], lam16[i]) for k in range(ntr)] for i in range(P)])
        return (float(r.time_ms[idx.max()]) if (idx >= 0).all() else None), idx

    with AmmBatch(probs, trials=16) as b:
        b.autotune_run(RunParams(step=2 * cfg["vr_period"], trial=20_000, **cfg), reps=2)
        dense_ms = b.time_dense(2)
        b.power_method(RunParams(step=iters, trial=10_000, **cfg))  # graph build + warm-up
        r = b.power_method(RunParams(step=iters, trial=0, **cfg))
        t16, idx16 = settle_all(r, 16)
        rx = b.power_method(RunParams(kernel="exact", step=args.exact_iters, trial=0))
        tx16, _ = settle_all(rx, 16)
    with AmmBatch(probs, trials=1) as b1:
        b1.autotune_run(RunParams(step=2 * cfg["vr_period"], trial=20_000, **cfg), reps=2)
        b1.power_method(RunParams(step=iters, trial=10_000, **cfg))
        r1 = b1.power_method(RunParams(step=iters, trial=0, **cfg))
        t1, _ = settle_all(r1, 1)
    gbs = 2 * elems * 2 / (dense_ms * 1e-3) / 1e9
    out = {
        "workload": f"{P} matrices, d~U[512,8192], n~U[1024,32768], s_i=10/i rank-64 + 0.01 N(0,1), BF16 "
                    "storage (leading dimension padded to 512 bytes); 16 seeds each",
        "elements": elems, "kernel": cfg["kernel"], "c": "n/8", "vr_period": cfg["vr_period"],
        "iterations": iters,
        "accounting": "per_solve = per (matrix, seed), 16 x matrices solves; per_matrix = run time / matrices",
        "batched16_time_all_ms": rnd(t16, 3), "settled_iteration_max": int(idx16.max()),
        "batched16_per_solve_ms": rnd(t16 / (16 * P) if t16 else None, 5),
        "batched16_per_matrix_ms": rnd(t16 / P if t16 else None, 4),
        "batched1_time_all_ms": rnd(t1, 3), "batched1_per_matrix_ms": rnd(t1 / P if t1 else None, 4),
        "exact16_bf16_time_all_ms": rnd(tx16, 3),
        "exact16_bf16_per_solve_ms": rnd(tx16 / (16 * P) if tx16 else None, 5),
        "exact_f32_total_ms": round(tot_exact, 3), "exact_f32_per_matrix_ms": round(tot_exact / P, 4),
        "exact_f32_vs_batched16_per_solve": ratio(tot_exact / P, t16 / (16 * P) if t16 else None, 2),
        "exact_f32_vs_batched16_per_matrix": ratio(tot_exact, t16, 3),
        "exact_f32_vs_batched1_per_matrix": ratio(tot_exact, t1, 3),
        "exact16_vs_batched16_same_engine": ratio(tx16, t16, 3),
        "final_rel_err_max": float(np.max(np.abs(r.eig[:, :, -1] - lam16[:, None]) / lam16[:, None])),
        "dense_pass_ms": round(dense_ms, 3), "dense_pass_GBs": round(gbs, 1),
        "dense_pass_frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 3),
    }
    del m16, probs
    torch.cuda.empty_cache()
    log(f"c4: {out}")
    return out


# ---------------------------------------------------------------------------
# C5: differentiable sigma_max of a BF16 weight stack
# ---------------------------------------------------------------------------
def c5_section(args):
    """sigma_max of a stack of 64 weight-like matrices W_k = G / sqrt(4096)
    + sum_{i<4} s_i u_i v_i^T (s = 5, 4, 3, 2; Haar u, v), 4096 x 4096 BF16:
    forward = one batched AMM power method (previousAMM c = 1024, exact step
    every 5, 50 iterations) + one exact pass for v, backward = the rank-1
    adjoint per matrix in one launch.  Checked against an FP64 power
    iteration on the stored values of two matrices; the FP32 exact forward
    of one matrix and the batched engine's exact forward of the whole stack
    beside it."""
    import torch
    from eigenvaluecomputation_by_amm_amd import RunParams
    from eigenvaluecomputation_by_amm_amd.autograd import sigma_max, sigma_max_stack
    from eigenvaluecomputation_by_amm_amd.workloads import _haar
    B, m = args.stack, 4096
    gen = torch.Generator(device="cuda").manual_seed(17)
    W = torch.empty(B, m, m, dtype=torch.bfloat16, device="cuda")
    sv = torch.tensor([5.0, 4.0, 3.0, 2.0], device="cuda", dtype=torch.float64)
    for k in range(B):
        G = torch.randn(m, m, generator=gen, device="cuda", dtype=torch.float32) / math.sqrt(m)
        U = _haar(m, 4, gen, "cuda")
        V = _haar(m, 4, gen, "cuda")
        W[k] = (G + ((U * sv) @ V.T).float()).to(torch.bfloat16)
    torch.cuda.synchronize()

    def fwd_bwd(params, tune, reps):
        Wg = W.requires_grad_(True)
        s = sigma_max_stack(Wg, params, tune=tune)  # warm-up: batch, tuned shapes, graph
        s.sum().backward()
        Wg.grad = None
        torch.cuda.synchronize()
        fw, bw = [], []
        for _ in range(reps):
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            s = sigma_max_stack(Wg, params)
            e1.record()
            s.sum().backward()
            e2.record()
            torch.cuda.synchronize()
            fw.append(e0.elapsed_time(e1))
            bw.append(e1.elapsed_time(e2))
            Wg.grad = None
        return statistics.median(fw), statistics.median(bw), Wg

    params = RunParams(kernel="previousAMM", c=1024, step=50, vr_period=5, seed=1)
    fwd, bwd, Wg = fwd_bwd(params, True, args.sigma_reps)
    xfwd, _, _ = fwd_bwd(RunParams(kernel="exact", step=50, seed=1), False, args.sigma_reps)
    s = sigma_max_stack(Wg, params)
    s.sum().backward()
    errs_s, errs_g = [], []
    for k in (0, B - 1):
        M = W[k].detach().double()
        x = torch.randn(m, generator=gen, device="cuda", dtype=torch.float64)
        for _ in range(300):
            x = M.T @ (M @ x)
            x /= torch.linalg.norm(x)
        t = M @ x
        sg = float(torch.linalg.norm(t))
        errs_s.append(abs(float(s[k].detach()) - sg) / sg)
        Gref = torch.outer(t / sg, x)
        errs_g.append(float((Wg.grad[k].double() - Gref).abs().max() / Gref.abs().max()))
    pe = RunParams(kernel="exact", step=50)
    W32 = W[0].detach().float().contiguous()
    sigma_max(W32, pe)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(4):
        sigma_max(W32, pe)
    e1.record()
    torch.cuda.synchronize()
    ex_one = e0.elapsed_time(e1) / 4
    Wg.grad = None
    del W, Wg, W32
    torch.cuda.empty_cache()
    out = {
        "workload": f"{B} x 4096x4096 BF16, G/64 + rank-4 spike (5,4,3,2)", "kernel": "previousAMM",
        "c": 1024, "vr_period": 5, "iterations": 50,
        "accounting": "one seed per matrix: per solve = per matrix",
        "forward_ms": round(fwd, 3), "backward_ms": round(bwd, 3),
        "per_matrix_ms": round((fwd + bwd) / B, 4),
        "exact_bf16_stack_forward_ms": round(xfwd, 3),
        "exact_f32_single_forward_ms": round(ex_one, 3),
        "exact_f32_stack_forward_ms_estimate": round(ex_one * B, 3),
        "exact_f32_vs_amm_forward": ratio(ex_one * B, fwd, 3),
        "exact_bf16_stack_vs_amm_forward_same_engine": ratio(xfwd, fwd, 3),
        "sigma_rel_err_vs_fp64_max": max(errs_s), "grad_rel_err_vs_fp64_max": max(errs_g),
    }
    log(f"c5: {out}")
    return out


def line(metric, value, unit, steps, warmup, ms_per_step, hib, dtype, config, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
         "ms_per_step": ms_per_step, "higher_is_better": hib, "scaling": "weak", "vs_baseline": None,
         "dtype": dtype, "data": "synthetic", "config": config}
    d.update(extra)
    return d


def probe_config(args):
    """Each hot kernel a few times on the 16384 x 65536 heavy-tailed
    workload, for hardware-counter profiling (`rocprofv3 --pmc ... --
    python3 bench.py --config probe`).  Launches, in order: a plain device
    copy of 1 GiB (read 1 GiB, write 1 GiB: the scale for FETCH_SIZE /
    WRITE_SIZE), the column-statistics pass, the statistics pass with
    z = A alpha, the one-pass exact product, the two-pass exact product (both
    contractions), one sampled product of the headline configuration and
    one batched 16-trial product in BF16 with every column present (both
    MFMA contractions); each `--probe-reps` times."""
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import heavy_tailed
    reps = args.probe_reps
    buf = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    out = torch.empty_like(buf)
    for _ in range(reps):
        out.copy_(buf)
    torch.cuda.synchronize()
    del buf, out
    A = heavy_tailed(args.d, args.n)
    with AmmContext(A) as ctx:
        ctx.time_column_stats(iters=reps)
        ctx.time_pass("stats_z", iters=reps)
        ctx.time_pass("exact", iters=reps)
        ctx.time_product("exact", 0, iters=reps)
        ctx.time_product(HEADLINE["kernel"], HEADLINE["c"], seed=1, iters=reps)
    torch.cuda.synchronize()
    Ab = A.to(torch.bfloat16)
    del A
    with AmmBatch([dict(A=Ab, c=HEADLINE["c"], seed=1)], trials=16) as b:
        b.power_method(RunParams(kernel="exact", step=reps))
    torch.cuda.synchronize()
    print(json.dumps({"probe": "done", "reps": reps, "workload": f"{args.d}x{args.n} heavy-tailed FP32"}),
          flush=True)


def probe_mix_config(args):
    """The batched contractions on the many-matrix mix (`--count` matrices,
    d ~ U[512, 8192], n ~ U[1024, 32768], BF16, leading dimension padded to
    512 bytes, N(0, 1) entries: traffic does not depend on the values), 16
    trials, for hardware-counter profiling.  Launches, in order: a plain
    device copy of 1 GiB (the scale for FETCH_SIZE / WRITE_SIZE), then
    `--probe-reps` dense passes of both contractions (every column present:
    k_bdot, k_bweight, k_baxpy)."""
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch
    from eigenvaluecomputation_by_amm_amd.workloads import padded_ld, uneven_problem_shapes
    reps = args.probe_reps
    buf = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    out = torch.empty_like(buf)
    for _ in range(reps):
        out.copy_(buf)
    torch.cuda.synchronize()
    del buf, out
    shapes = uneven_problem_shapes(args.count)
    mats = []
    for i, (d, n) in enumerate(shapes):
        g = torch.Generator(device="cuda").manual_seed(1000 + i)
        M = torch.zeros(n, padded_ld(d), dtype=torch.bfloat16, device="cuda")
        M[:, :d] = torch.randn(n, d, generator=g, device="cuda").to(torch.bfloat16)
        mats.append(M)
    with AmmBatch([dict(A=m, d=d, c=max(1, n // 8), seed=77 + i) for i, (m, (d, n)) in enumerate(zip(mats, shapes))],
                  trials=16) as b:
        ms = b.time_dense(reps)
    elems = sum(d * n for d, n in shapes)
    print(json.dumps({"probe": "done", "reps": reps, "workload": f"{len(shapes)} uneven BF16 matrices",
                      "elements": elems, "dense_pass_ms": round(ms, 3),
                      "algorithmic_bytes_per_contraction": elems * 2}), flush=True)


def kernel_roofline(ctx, A, d, n, n_exact, n_amm, kernel_iters):
    """Per-kernel durations (HIP events on the context stream) and each
    kernel's share of the headline run up to tolerance; the roofline entry
    is the kernel with the largest share."""
    import torch
    kname = "proposedAMM" if HEADLINE["kernel"] != "previousAMM" else "previousAMM"
    ms_dot, ms_axpy, ucols = ctx.time_product(kname, HEADLINE["c"], seed=1, iters=kernel_iters)
    ms_stats = ctx.time_column_stats(iters=10)
    s_A = A.element_size()
    kern = {}
    if d <= GRAM_MAX_ROWS:
        ms_gram = ctx.time_pass("exact", iters=20)
        kern[f"k_gram<float,{gram_L(d)}> (exact product A A^T x, one read of A)"] = (
            n_exact * ms_gram, ms_gram, d * n * s_A)
        ex_dot, ex_axpy, _ = ctx.time_product("exact", 0, iters=20)
        two_pass = {"k_gather_dot<float,4,dense> (two-pass exact, A^T x)": round(ex_dot * 1e3, 2),
                    "k_gather_axpy<float,dense> (two-pass exact, A t)": round(ex_axpy * 1e3, 2)}
    else:
        ex_dot, ex_axpy, _ = ctx.time_product("exact", 0, iters=20)
        kern["k_gather_dot<float,4,dense> (exact pass A^T x)"] = (n_exact * ex_dot, ex_dot, d * n * s_A)
        kern["k_gather_axpy<float,dense> (exact pass A t)"] = (n_exact * ex_axpy, ex_axpy, d * n * s_A)
        two_pass = {}
    kern.update({
        "k_colstats<float> (column statistics)": (ms_stats, ms_stats, d * n * s_A),
        "k_gather_dot<float,2,4,list> (sampled pass A_S^T dx)": (n_amm * ms_dot, ms_dot, ucols * d * s_A),
        "k_axpy_finish<float> (sampled pass A_S t with the finish)": (n_amm * ms_axpy, ms_axpy, ucols * d * s_A),
    })
    for k, (tot, t1, b) in kern.items():
        log(f"kernel {k}: {t1*1e3:.1f} us/launch, {b/(t1*1e-3)/1e9:.0f} GB/s, {tot*1e3:.0f} us in the run")
    dom_name, (dom_tot, dom_t, dom_b) = max(kern.items(), key=lambda kv: kv[1][0])
    achieved = dom_b / (dom_t * 1e-3) / 1e9
    buf = torch.empty(1 << 29, dtype=torch.float32, device="cuda")
    out = torch.empty_like(buf)
    out.copy_(buf)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        out.copy_(buf)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 10 * 2 * buf.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del buf, out
    read_ms = ctx.time_pass("read", iters=10)
    read_gbs = d * n * 4 / (read_ms * 1e-3) / 1e9
    log(f"torch device copy {copy_gbs:.0f} GB/s; streaming read of A {read_ms * 1e3:.1f} us, {read_gbs:.0f} GB/s")
    traffic = PMC_TRAFFIC_16384x65536.get(dom_name) if (d, n) == (16384, 65536) else None
    # the sampled iteration as a whole: both passes' bytes over the run's own
    # per-iteration device time is reported by the caller
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=dom_name,
                bytes_per_launch=int(dom_b), avg_launch_us=round(dom_t * 1e3, 2),
                share_of_run_us=round(dom_tot * 1e3, 1),
                kernels_us={k: round(v[1] * 1e3, 2) for k, v in kern.items()},
                not_in_run_us=two_pass, sampled_columns=ucols, torch_copy_GBs=round(copy_gbs, 1),
                read_ceiling_GBs=round(read_gbs, 1), frac_of_read_ceiling=round(achieved / read_gbs, 4))


def sampled_iteration(ctx, RunParams, lam, d, iters, ucols, s_A=4):
    """Device time of one sampled iteration of the headline configuration
    (run stamps: the median gap between consecutive sampled iterations) and
    its algorithmic bytes 2 u d s_A over it, against 8 TB/s."""
    r = ctx.power_method(RunParams(step=iters, seed=1, **HEADLINE))
    modes = ctx.last_modes(iters)
    gaps = [r.time_ms[i] - r.time_ms[i - 1] for i in range(1, iters) if modes[i] in (2, 3)]
    us = statistics.median(gaps) * 1e3 if gaps else None
    b = 2 * ucols * d * s_A
    return dict(us=rnd(us, 2), bytes=int(b), GBs=rnd(b / (us * 1e-6) / 1e9 if us else None, 1),
                frac_of_8TBs=rnd(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS if us else None, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed AMM runs (seeds)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--iterations", type=int, default=200)
    ap.add_argument("--exact-seeds", type=int, default=20)
    ap.add_argument("--table-seeds", type=int, default=20, help="seeds per cell of the c2 table")
    ap.add_argument("--sweep", action="store_true", help="also sweep kernel x c x vr (stderr table)")
    ap.add_argument("--sweep-seeds", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the c3, c4 and c5 sections")
    ap.add_argument("--d", type=int, default=16384)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--kernel-iters", type=int, default=200)
    ap.add_argument("--batch-reps", type=int, default=3, help="batched runs (16 seeds each)")
    ap.add_argument("--probe-reps", type=int, default=3, help="launches per kernel of --config probe")
    ap.add_argument("--config", choices=["headline", "batch", "bf16", "sigma", "seeds", "probe", "probe-mix"],
                    default="headline",
                    help="batch: many uneven matrices x 16 seeds; bf16: image-like BF16 workload; "
                         "sigma: weight stack; seeds: distribution over --seeds seeds of every "
                         "kernel x c on the headline matrix; probe: kernels for counter profiling "
                         "(separate JSON lines); probe-mix: the batched contractions on the "
                         "many-matrix mix, for counter profiling")
    ap.add_argument("--seeds", type=int, default=100, help="seeds per cell of --config seeds")
    ap.add_argument("--count", type=int, default=256, help="matrices of the c4 section")
    ap.add_argument("--batch-iters", type=int, default=150)
    ap.add_argument("--batch-vr", type=int, default=3,
                    help="variance-reduction period of the c4 section (3: fastest of 2-5 measured)")
    ap.add_argument("--exact-iters", type=int, default=40)
    ap.add_argument("--stack", type=int, default=64, help="matrices of the c5 section")
    ap.add_argument("--sigma-reps", type=int, default=3)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the batched form would shard by matrix, a single matrix by column blocks)", file=sys.stderr)
        sys.exit(2)

    if args.config == "probe":
        probe_config(args)
        return
    if args.config == "probe-mix":
        probe_mix_config(args)
        return
    if args.config == "batch":
        t0 = time.perf_counter()
        c4 = c4_section(args)
        print(json.dumps(line("lambda_max solves per second within 1e-3 rel err (batched BF16 AMM, 256 "
                              "matrices x 16 seeds), 1 MI355X",
                              rnd(16 * args.count / (c4["batched16_time_all_ms"] * 1e-3), 1)
                              if c4["batched16_time_all_ms"] else None, "solves/s", 1, 1,
                              round((time.perf_counter() - t0) * 1e3, 1), True, "bf16", c4)), flush=True)
        return
    if args.config == "bf16":
        c3 = c3_section(args, reps=args.batch_reps)
        print(json.dumps(line("time to lambda_max within 1e-3 rel err, BF16 storage (16 seeds per run), "
                              "1 MI355X", c3["batched16_per_matrix_ms"], "ms per matrix", args.batch_reps, 1,
                              c3["batched16_per_matrix_ms"], False, "bf16", c3)), flush=True)
        return
    if args.config == "sigma":
        c5 = c5_section(args)
        print(json.dumps(line("sigma_max forward + backward of a weight stack, 1 MI355X",
                              round(c5["forward_ms"] + c5["backward_ms"], 3), "ms", args.sigma_reps, 1,
                              round(c5["forward_ms"] + c5["backward_ms"], 3), False, "bf16", c5)), flush=True)
        return

    import torch
    from eigenvaluecomputation_by_amm_amd import AmmContext, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import heavy_tailed, lambda_max_lanczos

    d, n = args.d, args.n
    log(f"generating {d}x{n} FP32 heavy-tailed matrix on {torch.cuda.get_device_name(0)}")
    A = heavy_tailed(d, n)
    torch.cuda.synchronize()
    lam, change = lambda_max_lanczos(A, d, k=64)
    log(f"lambda* = {lam:.12e} (FP64 Lanczos change over last 8 steps {change:.1e})")
    ctx = AmmContext(A)
    iters = args.iterations
    if args.config == "seeds":
        t0 = time.perf_counter()
        cells = c2_table(ctx, RunParams, lam, list(range(1000, 1000 + args.seeds)), iters)
        ex = summarize(run_config(ctx, RunParams, lam, list(range(1000, 1010)), iters, kernel="exact"))
        print(json.dumps(line("distribution over seeds of the time to lambda_max within 1e-3 rel err, 1 MI355X",
                              cells["previousAMM c=1024"]["median_ms"], "ms", args.seeds, 0,
                              round((time.perf_counter() - t0) * 1e3 / max(1, 9 * args.seeds), 3), False, "f32",
                              {"workload": f"{d}x{n} FP32 heavy-tailed columns (lognormal 1.5) + planted rank 8",
                               "iterations": iters, "vr_period": 4, "tolerance": TOL, "lambda_star": lam,
                               "exact_10_seeds": ex, "cells": cells})), flush=True)
        return

    log("exact power method")
    ex = summarize(run_config(ctx, RunParams, lam, list(range(100, 100 + args.exact_seeds)), iters,
                              kernel="exact"))
    exact_rows = run_config(ctx, RunParams, lam, [100], iters, kernel="exact")
    log(f"exact: {ex}")

    if args.sweep:
        log("sweep kernel x c x vr")
        for kern in ["previousAMM", "proposedAMM", "proposedAMM_v2"]:
            for c in [256, 512, 1024, 2048, 4096, 16384]:
                for vr in [0, 2, 3, 4, 5, 10]:
                    if kern == "proposedAMM_v2" and vr == 0:
                        continue
                    if c in (256, 512, 2048) and vr in (0, 10):
                        continue
                    s = summarize(run_config(ctx, RunParams, lam, list(range(500, 500 + args.sweep_seeds)),
                                             iters, kernel=kern, c=c, vr_period=vr))
                    log(f"sweep {kern:15s} c={c:6d} vr={vr:3d}: median {s['median_ms']} ms iters "
                        f"{s['median_iters']} not_reached {s['not_reached']:.2f}")

    cfg = dict(HEADLINE)
    log(f"headline {cfg}: warmup {args.warmup}")
    run_config(ctx, RunParams, lam, list(range(900, 900 + args.warmup)), iters, **cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = run_config(ctx, RunParams, lam, list(range(1, 1 + args.steps)), iters, **cfg)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    am = summarize(rows)
    log(f"headline: {am}")
    # the same runs with the column statistics computed beforehand, as the
    # reference computes its norms while writing the matrix (get_norms in
    # device_memory.h:328-333), outside its clock
    keep = summarize(run_config(ctx, RunParams, lam, list(range(1, 1 + args.steps)), iters,
                                keep_stats=True, **cfg))

    log("c2 table: three kernels x c in {1024, 4096, 16384}")
    table = c2_table(ctx, RunParams, lam, list(range(1000, 1000 + args.table_seeds)), iters)

    ctx.power_method(RunParams(step=iters, seed=1, **cfg))  # the step kinds of the headline run
    modes = ctx.last_modes(iters)
    i_set = int(am["median_iters"]) if am["median_iters"] is not None else iters - 1
    n_exact = int(np.sum((modes[: i_set + 1] == 1) | (modes[: i_set + 1] == 4)))
    n_amm = i_set + 1 - n_exact
    roof = kernel_roofline(ctx, A, d, n, n_exact, n_amm, args.kernel_iters)
    roof["sampled_iteration"] = sampled_iteration(ctx, RunParams, lam, d, iters, roof["sampled_columns"])

    cpu = None
    if not args.no_cpu and am["median_ms"] is not None and rows[0]["i_settled"] >= 0:
        log("cpu baseline")
        Ah = A.cpu().numpy()
        cpu = cpu_baseline(Ah, d, lam, dict(cfg), rows[0]["seed"], rows[0]["i_settled"], exact_rows[0]["seed"],
                           exact_rows[0]["i_settled"], args)
        del Ah
        log(f"cpu: {cpu}")

    log("batched engine on the same matrix (16 seeds per run)")
    batched = batched_trials(A, d, n, lam, iters, dict(cfg), args.batch_reps, ex["median_ms"], am["median_ms"])
    ctx.close()
    del A
    torch.cuda.empty_cache()
    sections = {}
    if not args.no_configs:
        log("c3 section: image-like BF16")
        sections["c3"] = c3_section(args)
        log("c4 section: 256 uneven matrices x 16 seeds")
        sections["c4"] = c4_section(args)
        log("c5 section: sigma_max of a BF16 weight stack")
        sections["c5"] = c5_section(args)

    value = am["median_ms"]
    config = {"workload": f"{d}x{n} FP32 heavy-tailed columns (lognormal 1.5) + planted rank 8",
              "kernel": cfg["kernel"], "c": cfg["c"], "vr_period": cfg["vr_period"],
              "iterations": iters, "tolerance": TOL, "lambda_star": lam,
              "amm_p10_ms": rnd(am["p10_ms"]), "amm_p90_ms": rnd(am["p90_ms"]), "amm_mean_ms": rnd(am["mean_ms"]),
              "amm_not_reached": am["not_reached"], "amm_median_iters": am["median_iters"],
              "amm_final_rel_err_median": am["final_rel_err_median"],
              "amm_final_rel_err_max": am["final_rel_err_max"],
              "exact_ms": rnd(ex["median_ms"]), "exact_not_reached": ex["not_reached"],
              "exact_seeds": args.exact_seeds, "exact_p10_ms": rnd(ex["p10_ms"]), "exact_p90_ms": rnd(ex["p90_ms"]),
              "exact_mean_ms": rnd(ex["mean_ms"]), "exact_final_rel_err_max": ex["final_rel_err_max"],
              "exact_median_iters": ex["median_iters"],
              "speedup_vs_exact": ratio(ex["median_ms"], value),
              "speedup_vs_exact_p10": ratio(ex["p10_ms"], value),
              "amm_stats_outside_clock_ms": rnd(keep["median_ms"]),
              "speedup_vs_exact_stats_outside_clock": ratio(ex["median_ms"], keep["median_ms"]),
              "c2_table": {"seeds_per_cell": args.table_seeds, "vr_period": 4, "cells": table}}
    out = line(METRIC, rnd(value), "ms", args.steps, args.warmup, round(wall * 1e3 / args.steps, 3), False,
               "f32", config, roofline=roof, cpu_baseline=cpu, batched=batched, **sections)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
