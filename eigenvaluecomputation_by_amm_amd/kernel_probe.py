"""Run each hot kernel of the power method a few times on the
16384 x 65536 heavy-tailed workload, for hardware-counter profiling
(`rocprofv3 --pmc ... -- python3 -m eigenvaluecomputation_by_amm_amd.kernel_probe`).

Launches, in order: a plain device copy of 1 GiB (read 1 GiB, write 1 GiB:
the scale for FETCH_SIZE / WRITE_SIZE), the column-statistics pass, the
statistics pass with z = A alpha, the one-pass exact product, the two-pass
exact product (both contractions), one sampled product of the headline
configuration (previousAMM, c = 512) and one batched 16-trial product in
BF16 with every column present (both MFMA contractions).  Each runs
`--reps` times.
"""
from __future__ import annotations

import argparse
import sys


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--d", type=int, default=16384)
    ap.add_argument("--n", type=int, default=65536)
    args = ap.parse_args()
    import torch
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    from eigenvaluecomputation_by_amm_amd.workloads import heavy_tailed

    buf = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    out = torch.empty_like(buf)
    for _ in range(args.reps):
        out.copy_(buf)
    torch.cuda.synchronize()
    del buf, out
    A = heavy_tailed(args.d, args.n)
    with AmmContext(A) as ctx:
        ctx.time_column_stats(iters=args.reps)
        ctx.time_pass("stats_z", iters=args.reps)
        ctx.time_pass("exact", iters=args.reps)
        ctx.time_product("exact", 0, iters=args.reps)
        ctx.time_product("previousAMM", 512, seed=1, iters=args.reps)
    torch.cuda.synchronize()
    Ab = A.to(torch.bfloat16)
    del A
    with AmmBatch([dict(A=Ab, c=1024, seed=1)], trials=16) as b:
        b.power_method(RunParams(kernel="exact", step=args.reps))
    torch.cuda.synchronize()
    print("kernel_probe done", file=sys.stderr)


if __name__ == "__main__":
    main()
