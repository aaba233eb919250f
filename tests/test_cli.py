"""The AMM command line (drop-in for the reference's AMM.cu): flags, stdout
CSV, result/*.binary records, and the synthetic symmetric generator."""
import ctypes as C
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
AMM = ROOT / "eigenvaluecomputation_by_amm_amd" / "AMM"


def run_amm(args, cwd, timeout=120):
    return subprocess.run([str(AMM), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


# --- argument handling (host_memory.h:65-187); exits before any GPU call ---
@pytest.mark.parametrize("args,msg", [
    (["-kernel", "foo"], "[error] '-kernel foo' is not valid."),
    (["-kernel", "previousAMM", "-c", "7"], "multiple of 8"),
    (["-kernel", "previousAMM", "-m", "128", "-c", "256"], "1 <= c <= m"),
    (["-kernel", "proposedAMM_v2", "-c", "64"], "proposedAMM_v2 needs 1 <= vr <= step"),
    (["-dtype", "fp8"], "'-dtype fp8' is not valid"),
])
def test_argument_errors(tmp_path, args, msg):
    r = run_amm(args, tmp_path)
    assert r.returncode == 1
    assert msg in r.stdout + r.stderr


def test_missing_file(tmp_path):
    r = run_amm(["-filepath", tmp_path / "nope.bin"], tmp_path)
    assert r.returncode == 1 and "Failed to open file." in r.stderr


def write_matrix_file(path: Path, A: np.ndarray):
    """-filepath format: uint32 n, then n rows of n float32 (host_memory.h:143-159)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<I", A.shape[0]))
        f.write(np.ascontiguousarray(A, np.float32).tobytes())


def parse_csv(out: str):
    lines = out.strip().splitlines()
    assert lines[0] == "Time,Eigenvalue,NormError"
    rows = [tuple(float(v) for v in ln.split(",")) for ln in lines[1:]]
    for ln in lines[1:]:
        for v in ln.split(","):
            assert len(v.split(".")[1]) == 12  # %.12lf
    return np.array(rows)


@pytest.mark.gpu
def test_file_input_exact_and_binary(tmp_path):
    print("\n[gpu-case] cli file input exact", flush=True)
    rng = np.random.default_rng(0)
    n = 200  # padded to 256 by the CLI, as the reference does
    G = rng.standard_normal((n, n))
    A = ((G + G.T) / 2).astype(np.float32)
    w, V = np.linalg.eigh(A.astype(np.float64))
    k = np.argmax(np.abs(w))
    write_matrix_file(tmp_path / "sym.bin", A)
    (tmp_path / "result").mkdir()
    (tmp_path / "data").mkdir()
    V[:, k].astype(np.float64).tofile(tmp_path / "data" / "exact_eigenvector")
    r = run_amm(["-filepath", "sym.bin", "-kernel", "exact", "-step", "300", "-r", "2", "-display",
                 "-seed", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    rows = parse_csv(r.stdout)
    assert rows.shape == (600, 3)
    assert abs(rows[-1, 1] - abs(w[k])) / abs(w[k]) < 1e-4
    assert rows[-1, 2] < 1e-4
    assert np.all(np.diff(rows[:300, 0]) >= 0)
    f = tmp_path / "result" / "eigenvalues_by_exact_sym.bin_r=2_step=300.binary"
    rec = np.fromfile(f, np.float64).reshape(-1, 3)
    assert rec.shape == (600, 3)
    assert np.allclose(rec, rows, rtol=1e-11, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_synthetic_sanity_mode_tracks_exact(tmp_path, dtype):
    """Shifted synthetic matrices converge slowly (all eigenvalues sit near
    the Gershgorin offset), so compare trajectories: -sanity with c = m makes
    every AMM product exact, and the same seed gives the same x0."""
    print(f"\n[gpu-case] cli synthetic {dtype}", flush=True)
    (tmp_path / "result").mkdir()
    common = ["-m", "512", "-step", "40", "-r", "2", "-display", "-seed", "11", "-dtype", dtype,
              "-matrixType", "uniform_-1.0_1.0"]
    ex = parse_csv(run_amm(["-kernel", "exact", *common], tmp_path).stdout)
    for kern in ["previousAMM", "proposedAMM"]:
        r = run_amm(["-kernel", kern, "-c", "512", "-sanity", *common], tmp_path)
        assert r.returncode == 0, r.stderr
        rows = parse_csv(r.stdout)
        # Eigenvalue = max(0, sqrt(eig) - delta) can be 0 early on: compare on the run's scale
        assert np.max(np.abs(rows[:, 1] - ex[:, 1])) < 1e-5 * np.max(ex[:, 1])
    r = run_amm(["-kernel", "proposedAMM_v2", "-c", "256", "-vr", "5", *common], tmp_path)
    assert r.returncode == 0 and np.all(np.isfinite(parse_csv(r.stdout)))
    names = sorted(p.name for p in (tmp_path / "result").iterdir())
    assert "eigenvalues_by_proposedAMM_v2_uniform_-1.0_1.0_r=2_step=40_c=256_vr=5.binary" in names
    assert "eigenvalues_by_previousAMM_uniform_-1.0_1.0_r=2_step=40_c=512_vr=0.binary" in names


@pytest.mark.gpu
def test_file_input_amm_reaches_exact(tmp_path):
    """A symmetric matrix with a spectral gap: every AMM kernel with -vr 5
    ends within 1e-3 of the exact power method."""
    print("\n[gpu-case] cli amm vs exact", flush=True)
    rng = np.random.default_rng(1)
    n = 384
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    w = 10.0 / np.arange(1, n + 1)
    A = ((Q * w) @ Q.T + 0.01 * (lambda G: (G + G.T) / 2)(rng.standard_normal((n, n))))
    write_matrix_file(tmp_path / "gap.bin", A.astype(np.float32))
    (tmp_path / "result").mkdir()
    common = ["-filepath", "gap.bin", "-step", "40", "-r", "1", "-display", "-seed", "5"]
    lam = parse_csv(run_amm(["-kernel", "exact", *common], tmp_path).stdout)[-1, 1]
    ref = np.max(np.abs(np.linalg.eigvalsh(A.astype(np.float32).astype(np.float64))))
    assert abs(lam - ref) / ref < 1e-5
    for kern in ["previousAMM", "proposedAMM", "proposedAMM_v2"]:
        rows = parse_csv(run_amm(["-kernel", kern, "-c", "64", "-vr", "5", *common], tmp_path).stdout)
        assert abs(rows[-1, 1] - lam) / lam < 1e-3, (kern, rows[-1, 1], lam)


@pytest.mark.gpu
def test_generator_symmetric_with_gershgorin_offset():
    print("\n[gpu-case] generator", flush=True)
    import torch
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    m = 300
    for dist, p1, p2 in [(0, 0.0, 1.0), (1, 0.0, 0.5), (2, -2.0, 3.0)]:
        A = torch.empty((m, m), device="cuda")
        delta = C.c_float()
        assert lib.amm_generate_symmetric(C.c_void_p(A.data_ptr()), m, m, dist, p1, p2, 42, None,
                                          C.byref(delta)) == 0
        a = A.cpu().numpy().astype(np.float64)
        assert np.array_equal(a, a.T)
        raw = a - np.eye(m) * delta.value
        off = np.max(np.maximum(0, np.abs(raw).sum(0) - 2 * np.abs(np.diag(raw))))
        assert abs(off - delta.value) <= 1e-5 * off
        assert np.linalg.eigvalsh(a)[0] >= -1e-3 * off  # shifted spectrum is non-negative
        if dist == 2:
            assert raw.min() >= -2.0 and raw.max() < 3.0
        if dist == 1:
            assert raw.min() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("dist,p1,p2", [(0, 0.0, 1.0), (0, 1.5, 0.5), (1, 0.0, 0.5), (1, -1.0, 0.25),
                                        (2, -2.0, 3.0), (2, 0.5, 0.75)])
def test_generator_entry_distribution(dist, p1, p2):
    """The entries of the generated matrix (upper triangle with the diagonal,
    before the offset) follow the law that -matrixType names with its two
    parameters, as curandGenerateNormal(mean p1, std p2),
    curandGenerateLogNormal(p1, p2) and uniform p1 + (p2 - p1) u do
    (device_memory.h:21-32,402-424,487-497): one-sample KS at m = 2048
    (2.1 M entries, alpha = 0.01) plus the first two moments, so a std read
    as a variance or a swapped pair of parameters fails."""
    print(f"\n[gpu-case] generator law dist={dist} p1={p1} p2={p2}", flush=True)
    import torch
    from scipy import stats
    from eigenvaluecomputation_by_amm_amd import _lib
    lib = _lib.load()
    m = 2048
    A = torch.empty((m, m), device="cuda")
    delta = C.c_float()
    assert lib.amm_generate_symmetric(C.c_void_p(A.data_ptr()), m, m, dist, p1, p2, 7 + dist, None,
                                      C.byref(delta)) == 0
    a = A.cpu().numpy()
    np.fill_diagonal(a, np.diag(a) - np.float32(delta.value))
    e = a[np.triu_indices(m)].astype(np.float64)
    if dist == 0:
        law = stats.norm(loc=p1, scale=p2)
    elif dist == 1:
        law = stats.lognorm(s=p2, scale=np.exp(p1))
    else:
        law = stats.uniform(loc=p1, scale=p2 - p1)
    ks = stats.kstest(e, law.cdf)
    assert ks.pvalue > 0.01, ks
    se = law.std() / np.sqrt(e.size)
    assert abs(e.mean() - law.mean()) < 5 * se
    assert abs(e.std() / law.std() - 1.0) < 0.01
    # the diagonal carries the Gershgorin offset on top of the same law
    assert delta.value > 0


@pytest.mark.gpu
def test_cli_last_vr_flag_decides(tmp_path):
    """-vr and -vr_adaptive each set the reference's variance_reduction, and
    the last one parsed wins (host_memory.h:108-120); every variance-reduction
    step is gated on it (powerMethod_by_previous.cuh:135-138).  So
    '-vr 5 -vr_adaptive 0' runs no exact step at all (the same trajectory
    as '-vr 0'), '-vr_adaptive 0 -vr 5' runs one every 5 steps, and
    '-vr_adaptive 0.01 -vr 0' none either."""
    print("\n[gpu-case] cli vr flag order", flush=True)
    rng = np.random.default_rng(2)
    n = 256
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * (10.0 / np.arange(1, n + 1))) @ Q.T
    write_matrix_file(tmp_path / "m.bin", A.astype(np.float32))
    (tmp_path / "result").mkdir()
    common = ["-filepath", "m.bin", "-kernel", "previousAMM", "-c", "32", "-step", "20", "-r", "1",
              "-display", "-seed", "9"]

    def eig(*flags):
        r = run_amm([*common, *flags], tmp_path)
        assert r.returncode == 0, r.stderr
        return parse_csv(r.stdout)[:, 1]

    none = eig("-vr", "0")
    assert np.array_equal(eig("-vr", "5", "-vr_adaptive", "0"), none)
    assert np.array_equal(eig("-vr_adaptive", "0.01", "-vr", "0"), none)
    periodic = eig("-vr_adaptive", "0", "-vr", "5")
    assert np.array_equal(periodic[:4], none[:4])      # iterations 0-3 are sampled in both
    assert not np.array_equal(periodic[4:], none[4:])  # iteration 4 is exact with -vr 5
