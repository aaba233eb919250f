"""Reference-workflow tools around the command line: the FP64 eigen-solve of
matrix.dat (calculate_largest_eigenvalue.py) with the eigenvector file it
lacks, the result/*.binary reader, and the eigenfaces-style covariance
input (eigenfaces.py:15-58).  CPU tests, plus GPU tests that run the CLI."""
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
AMM = ROOT / "eigenvaluecomputation_by_amm_amd" / "AMM"


def _tools():
    from eigenvaluecomputation_by_amm_amd import covariance, eigh_check, results
    return eigh_check, results, covariance


def test_eigh_check_matrix_dat(tmp_path, capsys):
    eigh_check, _, _ = _tools()
    rng = np.random.default_rng(1)
    B = rng.standard_normal((40, 40))
    S = (B + B.T).astype(np.float32)
    S.tofile(tmp_path / "matrix.dat")
    out = tmp_path / "data" / "exact_eigenvector"
    eigh_check.main([str(tmp_path / "matrix.dat"), "--write-eigenvector", str(out)])
    lines = capsys.readouterr().out.strip().splitlines()
    w, V = np.linalg.eigh(S.astype(np.float64))
    assert lines[0] == "Largest eigenvalue  = {0:20f}".format(w.max())
    assert lines[1] == "Smallest eigenvalue = {0:20f}".format(w.min())
    v = np.fromfile(out, dtype=np.float64)
    k = int(np.argmax(np.abs(w)))
    assert v.size == 40 and abs(abs(v @ V[:, k]) - 1.0) < 1e-12
    # the vector is an eigenvector of the largest-magnitude eigenvalue
    assert np.allclose(S.astype(np.float64) @ v, w[k] * v, atol=1e-9)


def test_eigh_check_filepath_format(tmp_path):
    eigh_check, _, _ = _tools()
    A = np.arange(9, dtype=np.float32).reshape(3, 3)
    with open(tmp_path / "m.bin", "wb") as f:
        f.write(struct.pack("<I", 3))
        f.write(A.tobytes())
    assert np.array_equal(eigh_check.read_matrix(filepath=str(tmp_path / "m.bin")), A)
    with open(tmp_path / "bad.bin", "wb") as f:
        f.write(struct.pack("<I", 3))
        f.write(A.tobytes()[:20])
    with pytest.raises(ValueError):
        eigh_check.read_matrix(filepath=str(tmp_path / "bad.bin"))


def test_results_time_to_tolerance(tmp_path):
    _, results, _ = _tools()
    step = 6
    exact = 10.0
    # rep 0 settles at step 3 (one early hit at step 1 does not count), rep 1 never
    eig = np.array([[8.0, 10.0, 9.0, 10.005, 9.999, 10.0],
                    [8.0, 9.0, 9.5, 9.8, 9.9, 9.95]])
    t = np.cumsum(np.full((2, step), 0.5), axis=1)
    rec = np.stack([t, eig, np.zeros_like(eig)], axis=-1)
    rec.astype(np.float64).tofile(tmp_path / "r.binary")
    runs = results.read_binary(str(tmp_path / "r.binary"), step)
    assert runs.shape == (2, step, 3)
    res = results.time_to_tolerance(runs, exact, 1e-3)
    assert res[0] == (3, 2.0) and res[1] == (-1, None)
    with pytest.raises(ValueError):
        results.read_binary(str(tmp_path / "r.binary"), 5)


def test_results_plot(tmp_path, capsys):
    _, results, _ = _tools()
    step = 5
    t = np.cumsum(np.full((3, step), 0.25), axis=1)
    eig = 10.0 + np.array([[1.0, 0.1, 0.01, 1e-4, 1e-5]] * 3)
    np.stack([t, eig, np.zeros_like(eig)], axis=-1).astype(np.float64).tofile(tmp_path / "r.binary")
    out = tmp_path / "r.png"
    results.main([str(tmp_path / "r.binary"), "--step", str(step), "--exact", "10.0", "--plot", str(out)])
    assert "settled 3/3" in capsys.readouterr().out
    assert out.read_bytes()[:8] == b"\x89PNG\r\n\x1a\n"


def test_covariance_file(tmp_path):
    _, _, cov = _tools()
    rng = np.random.default_rng(2)
    imgs = rng.random((30, 6, 5)).astype(np.float32)
    np.save(tmp_path / "imgs.npy", imgs)
    out = tmp_path / "data" / "cov"
    cov.main([str(tmp_path / "imgs.npy"), "--out", str(out)])
    raw = open(out, "rb").read()
    (n,) = struct.unpack("<I", raw[:4])
    C = np.frombuffer(raw[4:], dtype=np.float32).reshape(n, n)
    X = imgs.reshape(30, -1).astype(np.float64).T            # pixels x images
    Xc = X - X.mean(axis=1, keepdims=True)                   # average face removed
    ref = Xc @ Xc.T / 30
    assert n == 30 and np.allclose(C, ref, rtol=1e-5, atol=1e-6)
    # lambda_max(C) = lambda_max(Xc Xc^T) / nimages, the rectangular identity --direct uses
    cols = cov.centred_columns(imgs)
    lam_rect = np.linalg.svd(cols.astype(np.float64), compute_uv=False)[0] ** 2 / 30
    assert abs(np.linalg.eigvalsh(ref)[-1] - lam_rect) < 1e-6 * lam_rect  # FP32 centring


def run_amm(args, cwd, timeout=120):
    return subprocess.run([str(AMM), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


@pytest.mark.gpu
def test_dump_matrix_eigh_and_normerror(tmp_path):
    """-dump_matrix writes matrix.dat (offset removed); its FP64 eigen-solve
    matches the exact power method's Eigenvalue column, and the eigenvector
    file it writes drives NormError to 0."""
    print("\n[gpu-case] cli dump_matrix + eigh_check", flush=True)
    eigh_check, _, _ = _tools()
    r = run_amm(["-m", "256", "-kernel", "exact", "-step", "300", "-r", "1", "-seed", "3",
                 "-matrixType", "uniform_0_1", "-display", "-dump_matrix"], tmp_path)
    assert r.returncode == 0, r.stderr
    A = eigh_check.read_matrix(str(tmp_path / "matrix.dat"))
    assert A.shape == (256, 256)
    w, _ = eigh_check.solve(A)
    eig = float(r.stdout.strip().splitlines()[-1].split(",")[1])
    assert abs(eig - w.max()) <= 1e-4 * abs(w.max())
    subprocess.run([sys.executable, "-m", "eigenvaluecomputation_by_amm_amd.eigh_check",
                    str(tmp_path / "matrix.dat"), "--write-eigenvector",
                    str(tmp_path / "data" / "exact_eigenvector")], check=True, cwd=ROOT)
    r2 = run_amm(["-m", "256", "-kernel", "exact", "-step", "300", "-r", "1", "-seed", "3",
                  "-matrixType", "uniform_0_1", "-display"], tmp_path)
    nerr = float(r2.stdout.strip().splitlines()[-1].split(",")[2])
    assert nerr < 1e-6


@pytest.mark.gpu
def test_covariance_direct_matches_file(tmp_path):
    """--direct (library on the rectangular centred images) agrees with the
    FP64 eigen-solve of the covariance file the CLI would read."""
    print("\n[gpu-case] covariance direct", flush=True)
    _, _, cov = _tools()
    from eigenvaluecomputation_by_amm_amd.workloads import grf_images
    imgs = grf_images(side=16, n=600, seed=3).cpu().numpy()
    Xc = cov.centred_columns(imgs)
    C = cov.covariance(Xc)
    lam = float(np.linalg.eigvalsh(C.astype(np.float64))[-1])
    assert abs(cov.lambda_direct(Xc, step=400) - lam) <= 1e-4 * lam
