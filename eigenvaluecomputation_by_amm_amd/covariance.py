"""Eigenfaces-style input from local images (eigenfaces.py:15-58).

The reference script downloads LFW, stacks the images as the columns of a
pixel x image matrix X, subtracts the average face from every column
(row-centring), forms the covariance C = Xc Xc^T / nimages in FP32 and
saves it in the command line's -filepath format (uint32 size, then size^2
float32).  Here the images come from a local .npy stack (nimages, h, w) or
(nimages, h*w); no download.

Two outputs:
  * the covariance file, for `AMM -filepath`, whose power method on C^2
    gives lambda_max(C);
  * with --direct, lambda_max(C) from the library on the rectangular Xc
    itself, no h*w x h*w covariance formed: lambda_max(C) =
    lambda_max(Xc Xc^T) / nimages.

    python -m eigenvaluecomputation_by_amm_amd.covariance images.npy \\
        --out data/covarianceMatrix_local [--direct]
"""
from __future__ import annotations

import argparse
import os
import struct

import numpy as np


def centred_columns(images: np.ndarray) -> np.ndarray:
    """(nimages, pixels) FP32, each pixel centred over the images: row j is
    column j (image j) of the centred pixel x image matrix Xc."""
    X = np.asarray(images, dtype=np.float64).reshape(images.shape[0], -1)
    return (X - X.mean(axis=0, keepdims=True)).astype(np.float32)


def covariance(Xc_cols: np.ndarray) -> np.ndarray:
    """C = Xc Xc^T / nimages (pixels x pixels), FP32 as the reference saves it."""
    Xc = Xc_cols.astype(np.float64)
    return (Xc.T @ Xc / Xc.shape[0]).astype(np.float32)


def write_filepath_format(path: str, C: np.ndarray) -> None:
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(struct.pack("<I", C.shape[0]))
        f.write(np.ascontiguousarray(C, np.float32).tobytes())


def lambda_direct(Xc_cols: np.ndarray, step: int = 200) -> float:
    """lambda_max(C) by the library's exact power method on Xc (device)."""
    import torch
    from .api import AmmContext, RunParams
    A = torch.from_numpy(np.ascontiguousarray(Xc_cols)).cuda()
    pad = (-A.shape[1]) % 4
    if pad:
        A = torch.nn.functional.pad(A, (0, pad))
    with AmmContext(A, d=Xc_cols.shape[1]) as ctx:
        r = ctx.power_method(RunParams(kernel="exact", step=step))
    return float(r.eig[-1]) / Xc_cols.shape[0]


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("images", help=".npy stack (nimages, h, w) or (nimages, pixels)")
    ap.add_argument("--out", default="data/covarianceMatrix_local")
    ap.add_argument("--direct", action="store_true")
    args = ap.parse_args(argv)
    imgs = np.load(args.images, allow_pickle=False)
    Xc = centred_columns(imgs)
    C = covariance(Xc)
    write_filepath_format(args.out, C)
    print(f"covariance {C.shape[0]} x {C.shape[0]} of {Xc.shape[0]} images -> {args.out}")
    if args.direct:
        print(f"lambda_max (library, on Xc) = {lambda_direct(Xc):.6f}")


if __name__ == "__main__":
    main()
