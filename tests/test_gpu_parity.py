"""GPU parity: the HIP library (through the C ABI) against the CPU oracle.

Bars: sampled indices, weights, CDF bit-exact; column statistics and
products within 1e-5 relative (FP32) of the FP64 oracle over the same
sample; power-method trajectories within 1e-4 relative of the oracle run on
the same samples; lambda_max against FP64 eigvalsh.
"""
import numpy as np
import pytest

from helpers import bf16_bits, bf16_to_f32, heavy_columns, lowrank_plus_noise, rel, to_cols
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _amm():
    from eigenvaluecomputation_by_amm_amd import AmmContext, RunParams, AmmError
    return AmmContext, RunParams, AmmError


def dev_matrix(A: np.ndarray, ld=None, bf16=False):
    """d x n numpy -> (tensor (n, ld) on cuda, host columns array)."""
    cols = to_cols(A, ld)
    if bf16:
        bits = bf16_bits(cols)
        t = torch.from_numpy(bits.view(np.int16)).cuda().view(torch.bfloat16)
        return t, bits
    return torch.from_numpy(cols).cuda(), cols


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


# --------------------------------------------------------------------------
# a wave per column under 16 KB, a workgroup per column from 16 KB (4100
# FP32 rows, 8200 BF16 rows), workgroups walking runs of columns for FP32
# columns of 64 KB or more (16384 x 2048: four columns each).  Short
# columns: whole trips of eight 64-lane steps (2048 FP32 / 4096 BF16 rows),
# whole steps of a partial trip, and the step that crosses d.
@pytest.mark.parametrize("d,n,ld,bf16", [(256, 300, 256, False), (1000, 777, 1004, False),
                                         (37, 5, 40, False), (512, 300, 512, True),
                                         (1001, 33, 1008, True), (2049, 64, 2052, False),
                                         (3000, 200, 3000, False), (4096, 100, 4096, True),
                                         (6000, 150, 6016, True), (4100, 300, 4104, False),
                                         (8200, 260, None, True), (16384, 2048, None, False)])
def test_column_stats(d, n, ld, bf16):
    banner(f"column_stats d={d} n={n} ld={ld} bf16={bf16}")
    AmmContext, *_ = _amm()
    A = lowrank_plus_noise(d, n, seed=1) + 0.25
    t, cols = dev_matrix(A, ld, bf16)
    with AmmContext(t, d=d) as ctx:
        alpha, nrm2, cnrm2, a2 = ctx.column_stats()
    ra, rn, rc = O.colstats(cols, d=d, bf16=bf16)
    assert rel(nrm2, rn) < 1e-12
    assert rel(cnrm2, rc) < 1e-9
    assert np.max(np.abs(alpha - ra) / np.maximum(np.abs(ra), 1e-30)) < 2e-7
    assert abs(a2 - float((alpha.astype(np.float64) ** 2).sum())) / a2 < 1e-12


# d = 16384 FP32: the statistics pass runs over runs of columns and hands
# its per-workgroup maxima to the CDF in place of the CDF's maximum pass
@pytest.mark.parametrize("d,n", [(128, 5000), (16384, 2048)])
@pytest.mark.parametrize("centred", [False, True])
def test_cdf_bitwise(centred, d, n):
    banner(f"cdf centred={centred} d={d} n={n}")
    AmmContext, *_ = _amm()
    A = heavy_columns(d, n, seed=2)
    A[:, 17] = 0.0
    t, _ = dev_matrix(A)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        cdf, tot, lp = ctx.build_cdf(centred)
    ocdf, otot, olp = O.cdf(cnrm2 if centred else nrm2)
    assert np.array_equal(cdf, ocdf) and tot == otot and lp == olp


@pytest.mark.parametrize("c,centred,iteration,trial", [(8, False, 0, 0), (1000, False, 3, 1),
                                                       (8192, True, 7, 0), (20000, False, 11, 5),
                                                       (4999, True, 2, 3)])
def test_sample_bitwise(c, centred, iteration, trial):
    banner(f"sample c={c} centred={centred} it={iteration} trial={trial}")
    AmmContext, *_ = _amm()
    A = heavy_columns(64, 25000, seed=3)
    A[:, [0, 100, 24999]] = 0.0
    t, _ = dev_matrix(A)
    seed = 0x1234_5678_9ABC
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        idx, wgt, segc = ctx.sample(c, iteration, seed, trial, centred)
    oi, ow, os_ = O.sample(cnrm2 if centred else nrm2, c, seed, iteration, trial)
    assert np.array_equal(segc, os_)
    for g, u in enumerate(segc):
        s = slice(g * O.SAMPLE_SEG, g * O.SAMPLE_SEG + int(u))
        assert np.array_equal(idx[s], oi[s])
        assert np.array_equal(wgt[s].view(np.uint32), ow[s].view(np.uint32))
    assert not np.isin([0, 100, 24999], idx[: segc[0]]).any()


@pytest.mark.parametrize("n", [65536, 90000, 3000])
def test_sample_bitwise_bucket_forms(n):
    """The sampler's search in its 8192-bucket coarse table, then within the
    bucket: buckets of exactly 8 columns (n = 65536), of 10-11 (n = 90000)
    and empty ones among single-column buckets (n = 3000 < 8192), each bit
    for bit against the oracle."""
    banner(f"sample bucket forms n={n}")
    AmmContext, *_ = _amm()
    A = heavy_columns(32, n, seed=41)
    A[:, [1, n - 2]] = 0.0
    t, _ = dev_matrix(A)
    seed = 0x5EED_0042
    with AmmContext(t) as ctx:
        _, nrm2, _, _ = ctx.column_stats()
        for c, it in ((3000, 0), (1024, 9)):
            idx, wgt, segc = ctx.sample(c, it, seed)
            oi, ow, os_ = O.sample(nrm2, c, seed, it)
            assert np.array_equal(segc, os_)
            u = int(segc[0])
            assert np.array_equal(idx[:u], oi[:u])
            assert np.array_equal(wgt[:u].view(np.uint32), ow[:u].view(np.uint32))
            assert not np.isin([1, n - 2], idx[:u]).any()


# rows per wave of the one-pass form: 1, 2, 4, 8, 16 loads per lane, and
# d = 16392 past it (the two-pass form); ragged d and padded ld included
@pytest.mark.parametrize("d,n,ld,bf16", [(512, 3000, None, False), (1000, 777, 1004, False),
                                         (1025, 300, 1028, False), (3001, 257, 3004, False),
                                         (8000, 130, None, False), (16384, 70, None, False),
                                         (16392, 50, None, False), (640, 900, None, True),
                                         (5000, 200, 5008, True), (16384, 60, None, True),
                                         (16392, 40, None, True), (1500, 1, None, False),
                                         (2049, 3, 2056, True), (7, 2, 8, False)])
def test_exact_product(d, n, ld, bf16):
    banner(f"exact_product d={d} n={n} ld={ld} bf16={bf16}")
    AmmContext, *_ = _amm()
    A = lowrank_plus_noise(d, n, seed=4)
    t, cols = dev_matrix(A, ld=ld, bf16=bf16)
    Ah = (bf16_to_f32(cols)[:, :d].T if bf16 else A).astype(np.float64)
    x = np.random.default_rng(5).standard_normal(d).astype(np.float32)
    with AmmContext(t, d=d) as ctx:
        y = ctx.exact_product(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel(y, Ah @ (Ah.T @ x)) < 1e-5


@pytest.mark.parametrize("kernel", ["previousAMM", "proposedAMM"])
@pytest.mark.parametrize("c,bf16", [(64, False), (1000, False), (9000, False), (256, True)])
def test_square_of_A(kernel, c, bf16):
    """One sampled product against the FP64 product over the same sample."""
    banner(f"square_of_A {kernel} c={c} bf16={bf16}")
    AmmContext, *_ = _amm()
    d, n = 520, 9500
    A = heavy_columns(d, n, seed=6) + 0.3
    t, cols = dev_matrix(A, bf16=bf16)
    colsf = bf16_to_f32(cols) if bf16 else cols
    rng = np.random.default_rng(7)
    x, xm, ym = (rng.standard_normal(d).astype(np.float32) for _ in range(3))
    seed, it = 99, 4
    prop = kernel == "proposedAMM"
    with AmmContext(t) as ctx:
        alpha, nrm2, cnrm2, a2 = ctx.column_stats()
        y = ctx.square_of_A(kernel, torch.from_numpy(x).cuda(), c, it, seed,
                            x_memo=torch.from_numpy(xm).cuda(), y_memo=torch.from_numpy(ym).cuda())
        y = y.cpu().numpy()
    idx, wgt, segc = O.sample(cnrm2 if prop else nrm2, c, seed, it)
    A64 = colsf.T.astype(np.float64)
    z = (A64 @ alpha.astype(np.float64)).astype(np.float32)
    v = (z.astype(np.float64) - a2).astype(np.float32)
    ref = O.amm_product(colsf, d, prop, c, idx, wgt, segc, False, x, xm, ym, alpha, v, z)
    assert rel(y, ref) < 1e-5


@pytest.mark.parametrize("kernel", ["previousAMM", "proposedAMM"])
def test_sanity_equals_exact(kernel):
    banner(f"sanity {kernel}")
    AmmContext, *_ = _amm()
    d, n = 300, 700
    A = lowrank_plus_noise(d, n, seed=8) + 1.0
    t, _ = dev_matrix(A)
    x = np.random.default_rng(9).standard_normal(d).astype(np.float32)
    with AmmContext(t) as ctx:
        y = ctx.square_of_A(kernel, torch.from_numpy(x).cuda(), n, 0, 1, sanity=True).cpu().numpy()
    A64 = A.astype(np.float64)
    assert rel(y, A64 @ (A64.T @ x)) < 1e-5


# --------------------------------------------------------------------------
RUNS = [
    dict(kernel="exact", step=30),
    dict(kernel="previousAMM", c=256, step=30),
    dict(kernel="previousAMM", c=256, step=30, vr_period=5),
    dict(kernel="proposedAMM", c=256, step=30, vr_period=4),
    dict(kernel="proposedAMM_v2", c=256, step=30, vr_period=5),
    dict(kernel="previousAMM", c=2048, step=12, sanity=True),
    dict(kernel="proposedAMM", c=2048, step=12, sanity=True),
    # variance reduction off with a period given (the reference's '-vr 5
    # -vr_adaptive 0'): no exact step, except proposedAMM_v2's transition
    dict(kernel="previousAMM", c=256, step=20, vr_period=5, no_vr=True),
    dict(kernel="proposedAMM", c=256, step=20, vr_period=5, no_vr=True),
    dict(kernel="proposedAMM_v2", c=256, step=20, vr_period=5, no_vr=True),
]


@pytest.mark.parametrize("run", RUNS, ids=lambda r: "-".join(f"{k}{v}" for k, v in r.items()))
def test_power_method_trajectory(run):
    """Whole runs against the oracle on the device's own column weights
    (same samples): per-iteration Rayleigh quotient within 1e-4."""
    banner(f"power_method {run}")
    AmmContext, RunParams, _ = _amm()
    d, n = 384, 2048
    A = lowrank_plus_noise(d, n, seed=10) + 0.05
    t, cols = dev_matrix(A)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=77, trial=2, **run))
        modes = ctx.last_modes(run["step"])
    ref = O.power_method(cols, d, seed=77, trial=2, w_plain=nrm2, w_cent=cnrm2, **run)
    omodes = ref["modes"].astype(np.uint32)
    omodes[omodes == 1] = 1
    # oracle codes 1 exact(+memo), 2 previous, 3 proposed; device adds 4 = exact without memo
    dm = modes.copy()
    dm[dm == 4] = 1
    assert np.array_equal(dm, omodes)
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4
    assert np.all(np.diff(res.time_ms) >= 0) and res.time_ms[0] > 0
    lam = O.lambda_max_eigvalsh(A)
    if run["kernel"] == "exact" or run.get("sanity"):
        assert abs(res.eig[-1] - lam) / lam < 1e-5


def test_adaptive_vr_matches_oracle():
    banner("adaptive vr")
    AmmContext, RunParams, _ = _amm()
    d, n = 256, 1024
    A = lowrank_plus_noise(d, n, seed=11)
    t, cols = dev_matrix(A)
    run = dict(kernel="previousAMM", c=128, step=25, vr_adaptive_threshold=0.02)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=5, **run))
        modes = ctx.last_modes(25)
    ref = O.power_method(cols, d, seed=5, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.array_equal(modes, ref["modes"].astype(np.uint32))
    assert (modes == 1).any() and (modes == 2).any()
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4


def test_graph_and_stream_launch_agree_bitwise():
    banner("graph vs eager")
    AmmContext, RunParams, _ = _amm()
    A = heavy_columns(256, 4000, seed=12)
    t, _ = dev_matrix(A)
    with AmmContext(t) as ctx:
        p = dict(kernel="proposedAMM_v2", c=512, step=20, vr_period=6, seed=3)
        a = ctx.power_method(RunParams(**p))
        b = ctx.power_method(RunParams(use_graph=False, **p))
        c = ctx.power_method(RunParams(**p))  # graph replay
    assert np.array_equal(a.eig, b.eig) and np.array_equal(a.eig, c.eig)


def test_concurrent_contexts_on_own_streams():
    """Two contexts on two streams driven from two host threads at once give
    the same bits as the same runs one after another: no state is shared
    between contexts beyond read-only launch shapes."""
    import threading
    banner("concurrent contexts")
    AmmContext, RunParams, _ = _amm()
    mats = [dev_matrix(heavy_columns(700, 5000, seed=30 + i))[0] for i in range(2)]
    runs = [dict(kernel="proposedAMM", c=600, step=25, vr_period=4, seed=5),
            dict(kernel="previousAMM", c=900, step=25, vr_period=5, seed=6)]

    def solve(i, out, stream):
        with AmmContext(mats[i], stream=stream) as ctx:
            out[i] = [ctx.power_method(RunParams(trial=t, **runs[i])).eig for t in range(3)]

    seq = [None, None]
    for i in range(2):
        solve(i, seq, 0)
    streams = [torch.cuda.Stream() for _ in range(2)]
    par = [None, None]
    th = [threading.Thread(target=solve, args=(i, par, streams[i].cuda_stream)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(2):
        assert par[i] is not None
        for a, b in zip(seq[i], par[i]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel", ["proposedAMM", "proposedAMM_v2"])
@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("d,ld", [(1000, 1000), (3001, 3008)])
def test_fused_z_equals_separate_pass(kernel, bf16, d, ld):
    """z = A alpha formed inside the column-statistics pass (a fresh run)
    against z from its own gather pass (a keep_stats run reuses the
    statistics, so only z's summation order differs): same samples, same
    modes, trajectories within 1e-5.  d = 3001 splits the rows unevenly
    over the waves."""
    banner(f"fused z {kernel} bf16={bf16} d={d}")
    AmmContext, RunParams, _ = _amm()
    A = lowrank_plus_noise(d, 3000, seed=13) + 0.2
    t, _ = dev_matrix(A, ld=ld, bf16=bf16)
    p = dict(kernel=kernel, c=400, step=20, vr_period=5, seed=9)
    with AmmContext(t, d=d) as ctx:
        a = ctx.power_method(RunParams(**p))
        ma = ctx.last_modes(20)
        b = ctx.power_method(RunParams(keep_stats=True, **p))
        mb = ctx.last_modes(20)
    assert np.array_equal(ma, mb)
    assert np.max(np.abs(a.eig - b.eig) / np.abs(b.eig)) < 1e-5


@pytest.mark.parametrize("d,n,ld", [(600, 1, 600), (9, 2, 12), (4100, 5, 4100)])
def test_fused_z_few_columns(d, n, ld):
    """Statistics + z with fewer columns than resident workgroups (one
    workgroup per column): proposed trajectories against the oracle."""
    banner(f"fused z d={d} n={n}")
    AmmContext, RunParams, _ = _amm()
    A = lowrank_plus_noise(d, n, seed=16) + 0.3
    t, cols = dev_matrix(A, ld=ld)
    run = dict(kernel="proposedAMM", c=n, step=6, vr_period=3)
    with AmmContext(t, d=d) as ctx:
        assert ctx.time_pass("stats_z", iters=1) > 0   # statistics + z in one pass
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=3, **run))  # z from the statistics pass
    ref = O.power_method(cols, d, seed=3, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4


@pytest.mark.parametrize("kernel", ["proposedAMM", "proposedAMM_v2"])
def test_bf16_fused_z_beyond_exact_form(kernel):
    """BF16 with 16384 < d <= 32768: z still rides in the statistics pass
    (16 loads per lane) while the exact step takes the two-pass form;
    trajectories against the oracle on the stored BF16 values."""
    banner(f"bf16 d=20000 {kernel}")
    AmmContext, RunParams, _ = _amm()
    d, n = 20000, 120
    A = lowrank_plus_noise(d, n, seed=17) + 0.2
    t, bits = dev_matrix(A, bf16=True)
    run = dict(kernel=kernel, c=60, step=8, vr_period=3)
    with AmmContext(t, d=d) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=4, **run))
    ref = O.power_method(bf16_to_f32(bits), d, seed=4, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4


def test_proposed_beyond_register_rows():
    """d above the register-held statistics form (16384 FP32 rows): the
    streamed statistics and a separate z pass, against the oracle."""
    banner("proposed d=16392")
    AmmContext, RunParams, _ = _amm()
    d, n = 16392, 300
    A = lowrank_plus_noise(d, n, seed=14) + 0.1
    t, cols = dev_matrix(A)
    run = dict(kernel="proposedAMM", c=64, step=8, vr_period=3)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=21, **run))
    ref = O.power_method(cols, d, seed=21, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4


def test_proposed_v2_beyond_register_rows():
    """proposedAMM_v2 in FP32 with d = 16392: z cannot ride in the
    statistics pass, so it is formed after the transition step's product
    ([z, tmp] = A [alpha, x], powerMethod_by_proposed_v2.cuh:255-294) and the
    transition's normalisation runs as its own kernel; against the oracle."""
    banner("proposed_v2 d=16392")
    AmmContext, RunParams, _ = _amm()
    d, n = 16392, 260
    A = lowrank_plus_noise(d, n, seed=18) + 0.1
    t, cols = dev_matrix(A)
    run = dict(kernel="proposedAMM_v2", c=64, step=10, vr_period=3)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=22, **run))
        modes = ctx.last_modes(10)
    ref = O.power_method(cols, d, seed=22, w_plain=nrm2, w_cent=cnrm2, **run)
    dm = modes.copy()
    dm[dm == 4] = 1
    assert np.array_equal(dm, ref["modes"].astype(np.uint32))
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4


@pytest.mark.parametrize("kernel", ["exact", "previousAMM", "proposedAMM", "proposedAMM_v2"])
def test_rows_beyond_fused_normalisation(kernel):
    """d = 70000 > 65536 rows: the normalisation x = y / ||y|| leaves the
    finish tail for a grid pass (and the proposed s, q dots for their own
    tail); every kernel against the oracle."""
    banner(f"d=70000 n=64 {kernel}")
    AmmContext, RunParams, _ = _amm()
    d, n = 70000, 64
    A = lowrank_plus_noise(d, n, seed=19) + 0.05
    t, cols = dev_matrix(A)
    run = dict(kernel=kernel, c=0 if kernel == "exact" else 32, step=8, vr_period=3)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=23, **run))
        x = ctx.get_vector("x")
    ref = O.power_method(cols, d, seed=23, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4
    assert abs(np.linalg.norm(x.astype(np.float64)) - 1.0) < 1e-5
    assert rel(x, ref["x"]) < 1e-3


def test_time_pass():
    """amm_time_pass times the statistics+z pass, the exact product as a run
    issues them and the plain read of A, and refuses the fused statistics
    past 16384 FP32 rows."""
    banner("time_pass")
    AmmContext, _, AmmError = _amm()
    for d, fused in ((2000, True), (16392, False)):
        A = lowrank_plus_noise(d, 64, seed=15)
        t, _ = dev_matrix(A)
        with AmmContext(t) as ctx:
            assert ctx.time_pass("exact", iters=3) > 0
            assert ctx.time_pass("read", iters=3) > 0   # the plain streaming read of A
            if fused:
                assert ctx.time_pass("stats_z", iters=3) > 0
            else:
                with pytest.raises(AmmError):
                    ctx.time_pass("stats_z", iters=3)


def test_lambda_small_workload_vs_eigvalsh():
    """The CPU-runnable workload (1024 x 2048, s_i = 10/i, c = 256, 100
    iterations): exact reaches the FP64 eigenvalue; previous AMM with VR
    reaches it within 1e-3."""
    banner("C1 lambda")
    AmmContext, RunParams, _ = _amm()
    A = lowrank_plus_noise(1024, 2048, seed=13)
    lam = O.lambda_max_eigvalsh(A)
    t, _ = dev_matrix(A)
    with AmmContext(t) as ctx:
        ex = ctx.power_method(RunParams(kernel="exact", step=100, seed=1))
        pv = ctx.power_method(RunParams(kernel="previousAMM", c=256, step=100, vr_period=10, seed=1))
    assert abs(ex.eig[-1] - lam) / lam < 1e-6
    assert abs(pv.eig[-1] - lam) / lam < 1e-3


def test_zero_matrix_and_zero_columns():
    banner("zero matrix")
    AmmContext, RunParams, _ = _amm()
    t = torch.zeros((300, 200), device="cuda")
    with AmmContext(t) as ctx:
        for k in ["exact", "previousAMM", "proposedAMM", "proposedAMM_v2"]:
            r = ctx.power_method(RunParams(kernel=k, c=64, step=6, vr_period=3))
            assert np.all(r.eig == 0.0) and np.all(np.isfinite(r.err))


def test_nonfinite_input_rejected():
    banner("nonfinite")
    AmmContext, RunParams, AmmError = _amm()
    A = lowrank_plus_noise(64, 100, seed=14)
    A[3, 7] = np.nan
    t, _ = dev_matrix(A)
    with AmmContext(t) as ctx:
        with pytest.raises(AmmError):
            ctx.power_method(RunParams(kernel="previousAMM", c=16, step=3))


def test_invalid_arguments():
    banner("invalid args")
    AmmContext, RunParams, AmmError = _amm()
    t = torch.ones((100, 64), device="cuda")
    with AmmContext(t) as ctx:
        with pytest.raises(AmmError):
            ctx.power_method(RunParams(kernel="previousAMM", c=101, step=3))
        with pytest.raises(AmmError):
            ctx.power_method(RunParams(kernel="proposedAMM_v2", c=8, step=3, vr_period=0))
        with pytest.raises(AmmError):
            ctx.power_method(RunParams(kernel="previousAMM", c=0, step=3))


def test_eigenvector_norm_error():
    banner("norm error")
    AmmContext, RunParams, _ = _amm()
    A = lowrank_plus_noise(200, 400, seed=15)
    w, V = np.linalg.eigh(A.astype(np.float64) @ A.T.astype(np.float64))
    t, _ = dev_matrix(A)
    with AmmContext(t) as ctx:
        ctx.set_exact_eigenvector(V[:, -1])
        r = ctx.power_method(RunParams(kernel="exact", step=60, seed=2))
    assert r.norm_error()[-1] < 1e-8
    assert abs(r.eigenvalue()[-1] - np.sqrt(w[-1])) / np.sqrt(w[-1]) < 1e-6


def test_clearing_eigenvector_drops_captured_graph():
    """A graph-launched run after amm_set_exact_eigenvector(ctx, NULL) must
    not keep reading the old vector: err is 0 again, as in an eager run."""
    banner("clear eigenvector, graph run")
    AmmContext, RunParams, _ = _amm()
    A = lowrank_plus_noise(200, 400, seed=15)
    w, V = np.linalg.eigh(A.astype(np.float64) @ A.T.astype(np.float64))
    t, _ = dev_matrix(A)
    p = RunParams(kernel="previousAMM", c=64, step=12, vr_period=4, seed=3)
    with AmmContext(t) as ctx:
        ctx.set_exact_eigenvector(V[:, -1])
        r1 = ctx.power_method(p)
        ctx.set_exact_eigenvector(None)
        r2 = ctx.power_method(p)
        r3 = ctx.power_method(RunParams(kernel="previousAMM", c=64, step=12, vr_period=4, seed=3,
                                        use_graph=False))
    assert np.abs(r1.err[-1]) > 0.5
    assert np.all(r2.err == 0.0) and np.all(r3.err == 0.0)
    assert np.array_equal(r2.eig, r3.eig)


@pytest.mark.parametrize("kernel", ["previousAMM", "proposedAMM"])
@pytest.mark.parametrize("c,bf16", [(700, False), (5000, False), (3000, True)])
def test_square_of_A_fused_finish(kernel, c, bf16):
    """Rows enough for the second contraction to carry the finish (d >= 8192,
    64-row blocks): one sampled product against the FP64 product over the
    same sample; c = 5000 spans three LDS chunks of the list."""
    banner(f"square_of_A fused finish {kernel} c={c} bf16={bf16}")
    AmmContext, *_ = _amm()
    d, n = 8200, 6000
    A = heavy_columns(d, n, seed=26) + 0.3
    t, cols = dev_matrix(A, bf16=bf16)
    colsf = bf16_to_f32(cols) if bf16 else cols
    rng = np.random.default_rng(27)
    x, xm, ym = (rng.standard_normal(d).astype(np.float32) for _ in range(3))
    seed, it = 31, 2
    prop = kernel == "proposedAMM"
    with AmmContext(t) as ctx:
        alpha, nrm2, cnrm2, a2 = ctx.column_stats()
        y = ctx.square_of_A(kernel, torch.from_numpy(x).cuda(), c, it, seed,
                            x_memo=torch.from_numpy(xm).cuda(), y_memo=torch.from_numpy(ym).cuda())
        y = y.cpu().numpy()
    idx, wgt, segc = O.sample(cnrm2 if prop else nrm2, c, seed, it)
    A64 = colsf.T.astype(np.float64)
    z = (A64 @ alpha.astype(np.float64)).astype(np.float32)
    v = (z.astype(np.float64) - a2).astype(np.float32)
    ref = O.amm_product(colsf, d, prop, c, idx, wgt, segc, False, x, xm, ym, alpha, v, z)
    assert rel(y, ref) < 1e-5


@pytest.mark.parametrize("run", [dict(kernel="previousAMM", c=600, step=20, vr_adaptive_threshold=0.05),
                                 dict(kernel="proposedAMM", c=600, step=20, vr_adaptive_threshold=0.05),
                                 dict(kernel="proposedAMM", c=3000, step=12, vr_period=4)],
                         ids=["adaptive-previous", "adaptive-proposed", "proposed-c3000"])
def test_fused_finish_runs_against_oracle(run):
    """Whole runs where sampled steps carry the finish in their second
    contraction: with adaptive variance reduction the finish kernel runs
    only on exact steps and the fused kernel only on sampled ones (both
    gated on the step kind the device decided).  Same samples as the
    oracle: identical step kinds, trajectories within 1e-4."""
    banner(f"fused finish run {run}")
    AmmContext, RunParams, _ = _amm()
    d, n = 8200, 4000
    A = lowrank_plus_noise(d, n, seed=28) + 0.05
    t, cols = dev_matrix(A)
    with AmmContext(t) as ctx:
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=29, **run))
        modes = ctx.last_modes(run["step"])
        eager = ctx.power_method(RunParams(seed=29, use_graph=False, **run))
    ref = O.power_method(cols, d, seed=29, w_plain=nrm2, w_cent=cnrm2, **run)
    assert np.array_equal(modes, ref["modes"].astype(np.uint32))
    if "vr_adaptive_threshold" in run:
        assert (modes == 1).any() and (modes != 1).any()
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4
    assert np.array_equal(res.eig, eager.eig)


@pytest.mark.parametrize("d", [8192, 4096])
@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("run", [dict(kernel="previousAMM", c=600, step=14, vr_period=4),
                                 dict(kernel="proposedAMM", c=600, step=14, vr_period=4),
                                 dict(kernel="proposedAMM_v2", c=600, step=14, vr_period=4)],
                         ids=["previous", "proposed", "proposed_v2"])
def test_carried_tail_runs_against_oracle(run, bf16, d):
    """Runs where each iteration's tail (lambda, err, 1/||y||, x = y / ||y||
    and the proposed s, q) rides in the next sampled dot pass: d is a
    multiple of the dot pass's row split (1024 FP32 / 2048 BF16 rows) and the
    schedule is static, so sampled -> sampled and exact -> sampled steps both
    carry it; d = 8192 with the finish in the second contraction, d = 4096
    with the chunked contraction and the finish kernel.  Same samples as the
    oracle: identical step kinds, Rayleigh quotients and err = y^T v* / ||y||
    within 1e-4; graph and eager runs bit for bit."""
    banner(f"carried tail run {run} bf16={bf16} d={d}")
    AmmContext, RunParams, _ = _amm()
    n = 3000
    A = lowrank_plus_noise(d, n, seed=30) + 0.05
    t, cols = dev_matrix(A, bf16=bf16)
    colsf = bf16_to_f32(cols) if bf16 else cols
    A64 = colsf.T.astype(np.float64)  # d x n
    v = np.random.default_rng(31).standard_normal(d)
    for _ in range(3):  # near the top left singular vector: err ~ 1, not ~ 0
        v = A64 @ (A64.T @ v)
        v /= np.linalg.norm(v)
    with AmmContext(t) as ctx:
        ctx.set_exact_eigenvector(v)
        _, nrm2, cnrm2, _ = ctx.column_stats()
        res = ctx.power_method(RunParams(seed=32, **run))
        modes = ctx.last_modes(run["step"])
        eager = ctx.power_method(RunParams(seed=32, use_graph=False, **run))
    ref = O.power_method(colsf, d, seed=32, w_plain=nrm2, w_cent=cnrm2, eigvec=v, **run)
    dm = modes.copy()
    dm[dm == 4] = 1  # the device's exact step without memo (v2 transition) is the oracle's 1
    assert np.array_equal(dm, ref["modes"].astype(np.uint32))
    assert ((modes == 2) | (modes == 3)).sum() >= 8  # sampled steps carry the tails
    assert np.max(np.abs(res.eig - ref["eig"]) / np.abs(ref["eig"])) < 1e-4
    assert np.max(np.abs(res.err - ref["err"])) < 1e-4
    assert np.abs(ref["err"][-1]) > 0.5  # err is not trivially zero
    assert np.array_equal(res.eig, eager.eig) and np.array_equal(res.err, eager.err)


def test_lanczos_ground_truth_against_eigvalsh():
    """The benchmark's FP64 ground truth (workloads.lambda_max_lanczos, GPU)
    against FP64 eigvalsh of A A^T (the computation of
    calculate_largest_eigenvalue.py:9-10) and against the oracle's numpy
    Lanczos, on the heavy-tailed construction at 2048 x 8192, FP32 and BF16."""
    banner("lanczos vs eigvalsh 2048x8192")
    from eigenvaluecomputation_by_amm_amd.workloads import heavy_tailed, lambda_max_lanczos
    A = heavy_tailed(2048, 8192, seed=5)
    for M in (A, A.to(torch.bfloat16)):
        lam, change = lambda_max_lanczos(M, 2048, k=64)
        M64 = M.double()
        lam_e = float(torch.linalg.eigvalsh(M64.T @ M64)[-1])
        Mh = M64.cpu().numpy()
        lam_o = O.lanczos_lambda_max(lambda v: Mh.T @ (Mh @ v), 2048, k=64)
        assert change < 1e-12
        assert abs(lam - lam_e) / lam_e < 1e-11, (lam, lam_e)
        assert abs(lam_o - lam_e) / lam_e < 1e-11, (lam_o, lam_e)


def test_statistical_parity_with_reference_sampler():
    """The CPU-size configuration (1024 x 2048, s_i = 10/i plus 0.01 N(0,1),
    previousAMM c = 256 with p_j ~ ||a_j||^2, no variance reduction, 100
    iterations) over 100 seeds: the distribution of the final relative
    error against FP64 eigvalsh matches that of the restatement of the
    reference's own sampler (per-slot xorshift32 states, 23-bit uniforms,
    FP32 CDF and weights, duplicates kept; device_memory.h:506-545,
    common.h:34-53, loop powerMethod_by_previous.cuh:132-180).  Each arm
    draws its own x0 ~ N(0, I) per seed.  Two-sample KS and k-sample
    Anderson-Darling, both at alpha = 0.01."""
    banner("statistical parity C1 100 seeds")
    from scipy import stats
    AmmContext, RunParams, _ = _amm()
    d, n, c, step = 1024, 2048, 256, 100
    A = lowrank_plus_noise(d, n, seed=13)
    lam = O.lambda_max_eigvalsh(A)
    t, cols = dev_matrix(A)
    ours, refs = [], []
    with AmmContext(t) as ctx:
        for s in range(100):
            r = ctx.power_method(RunParams(kernel="previousAMM", c=c, step=step, seed=1000 + s))
            ours.append(abs(r.eig[-1] - lam) / lam)
    for s in range(100):
        x0 = np.random.default_rng(5000 + s).standard_normal(d).astype(np.float32)
        x0 /= np.linalg.norm(x0)
        e = O.ref_power_method(cols, d, "previousAMM", c, step, 9000 + s, x0)
        refs.append(abs(e[-1] - lam) / lam)
    ours, refs = np.array(ours), np.array(refs)
    print(f"[gpu-case] final rel err median: library {np.median(ours):.3e}, reference sampler "
          f"{np.median(refs):.3e}", flush=True)
    assert stats.ks_2samp(ours, refs).pvalue > 0.01
    assert stats.anderson_ksamp([ours, refs]).significance_level > 0.01
