"""Read the command line's result/*.binary files and extract time to
tolerance.

A result file holds, for each repetition r and step i, three doubles
(cumulative time in ms, Eigenvalue, NormError) as printed in the CSV
(powerMethod_by_exact.cuh:71-99 and its siblings).  Eigenvalue is
max(0, sqrt(lambda_i) - delta), the square root of the Rayleigh quotient
of A A^T.

    python -m eigenvaluecomputation_by_amm_amd.results FILE.binary --step S
        (--exact VALUE | --matrix matrix.dat | --filepath F) [--tol 1e-3]

prints, per repetition, the first step after which |Eigenvalue - exact| <=
tol * exact holds for good and its time, then the median.  With --plot
PATH.png it also draws the relative error of every repetition against
time (log scale), with the tolerance as a line.
"""
from __future__ import annotations

import argparse
import statistics

import numpy as np


def read_binary(path: str, step: int) -> np.ndarray:
    """(repetitions, step, 3) array of (time_ms, eigenvalue, norm_error)."""
    data = np.fromfile(path, dtype=np.float64)
    if data.size % (3 * step):
        raise ValueError(f"{path}: {data.size} doubles is not a whole number of {step}-step runs")
    return data.reshape(-1, step, 3)


def settled(values: np.ndarray, exact: float, tol: float) -> int:
    """First i with |values[j] - exact| <= tol*|exact| for every j >= i
    (-1 if the last value is outside)."""
    ok = np.abs(values - exact) <= tol * abs(exact)
    if not ok[-1]:
        return -1
    bad = np.nonzero(~ok)[0]
    return 0 if bad.size == 0 else int(bad[-1] + 1)


def time_to_tolerance(runs: np.ndarray, exact: float, tol: float = 1e-3):
    out = []
    for r in runs:
        i = settled(r[:, 1], exact, tol)
        out.append((i, float(r[i, 0]) if i >= 0 else None))
    return out


def plot(runs: np.ndarray, exact: float, tol: float, path: str, label: str = "") -> None:
    """Relative error |Eigenvalue - exact| / |exact| of every repetition
    against cumulative time, log scale, tolerance as a dashed line."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, ax = plt.subplots(figsize=(6.4, 4.0))
    for r, run in enumerate(runs):
        rel = np.abs(run[:, 1] - exact) / abs(exact)
        ax.semilogy(run[:, 0], np.maximum(rel, 1e-17), lw=0.8, alpha=0.6,
                    color="C0", label=(label or "repetitions") if r == 0 else None)
    ax.axhline(tol, color="k", ls="--", lw=0.8, label=f"tolerance {tol:g}")
    ax.set_xlabel("time (ms)")
    ax.set_ylabel("relative error of Eigenvalue")
    ax.legend()
    fig.tight_layout()
    fig.savefig(path, dpi=120)
    plt.close(fig)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("binary")
    ap.add_argument("--step", type=int, required=True)
    ap.add_argument("--exact", type=float, default=None, help="reference Eigenvalue (sqrt of lambda_max)")
    ap.add_argument("--matrix", default=None, help="matrix.dat to solve for the reference")
    ap.add_argument("--filepath", default=None, help="-filepath matrix to solve for the reference")
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--plot", default=None, metavar="PATH.png")
    args = ap.parse_args(argv)
    exact = args.exact
    if exact is None:
        from .eigh_check import read_matrix, solve
        w, _ = solve(read_matrix(args.matrix, args.filepath))
        exact = float(np.max(np.abs(w)))
    runs = read_binary(args.binary, args.step)
    res = time_to_tolerance(runs, exact, args.tol)
    print("repetition,settled_step,time_ms")
    for r, (i, t) in enumerate(res):
        print(f"{r},{i},{'' if t is None else f'{t:.6f}'}")
    ts = [t for _, t in res if t is not None]
    print(f"# exact {exact:.12f}; settled {len(ts)}/{len(res)}; "
          f"median time {statistics.median(ts) if ts else float('nan'):.6f} ms")
    if args.plot:
        plot(runs, exact, args.tol, args.plot, label=args.binary.rsplit("/", 1)[-1])


if __name__ == "__main__":
    main()
