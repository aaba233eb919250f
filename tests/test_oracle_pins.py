"""Pins of the CPU oracle (oracle/amm_oracle.c) against published known
answers, closed forms, brute force and the laws of the estimator.  CPU only.

Each test is chosen so that a plausible slip in the oracle (wrong constant,
off-by-one search, dropped weight factor, transposed operand, wrong sign in
the mean shift) fails it.
"""
import numpy as np
import pytest
from scipy import stats

from helpers import heavy_columns, lowrank_plus_noise, rel, to_cols, bf16_bits, bf16_to_f32
from oracle import oracle as O


# --- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11) ---
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert list(O.philox4x32_10(ctr, key)) == expect


def test_uniform53_range_and_law():
    assert O.uniform53(0, 0) == 0.0
    assert O.uniform53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert O.uniform53(1 << 5, 0) == 2.0 ** -27
    assert O.uniform53(0, 1 << 6) == 2.0 ** -53
    # the Philox stream through uniform53 is uniform (KS)
    u = []
    for m in range(4000):
        r = O.philox4x32_10([m, 7, 0, 1], [12345, 0])
        u += [O.uniform53(int(r[0]), int(r[1])), O.uniform53(int(r[2]), int(r[3]))]
    assert stats.kstest(u, "uniform").pvalue > 1e-4


def test_x0_is_gaussian_and_unit():
    g = O.x0(seed=99, trial=3, d=100001, normalise=False).astype(np.float64)
    assert stats.kstest(g, "norm").pvalue > 1e-4
    x = O.x0(seed=99, trial=3, d=100001)
    assert abs(np.linalg.norm(x.astype(np.float64)) - 1.0) < 1e-6
    # x is g / ||g||
    assert rel(x, g / np.linalg.norm(g)) < 1e-6
    # odd length: last element comes from the cosine branch
    assert O.x0(5, 0, 7, normalise=False)[6] == O.x0(5, 0, 8, normalise=False)[6]


# --- xorshift32 (13,17,5) of the reference's get_rand (device_memory.h:506-514) ---
def test_xorshift32_kat():
    # Marsaglia, "Xorshift RNGs" (J. Stat. Softw. 2003): y = 2463534242 -> 723471715
    assert O.xorshift32(2463534242) == 723471715


def _xs_apply(v):
    v ^= (v << 13) & 0xFFFFFFFF
    v ^= v >> 17
    v ^= (v << 5) & 0xFFFFFFFF
    return v


def test_xorshift32_full_period_gf2():
    """(13,17,5) has period 2^32 - 1: the GF(2) matrix T satisfies
    T^(2^32-1) = I and T^((2^32-1)/p) != I for each prime p | 2^32-1."""
    cols = [_xs_apply(1 << i) for i in range(32)]
    # the oracle's step is the same linear map
    for i in range(32):
        assert O.xorshift32(1 << i) == cols[i]

    def apply(M, v):
        r = 0
        i = 0
        while v:
            if v & 1:
                r ^= M[i]
            v >>= 1
            i += 1
        return r

    def mul(A, B):
        return [apply(A, B[j]) for j in range(32)]

    def power(M, e):
        R = [1 << i for i in range(32)]
        while e:
            if e & 1:
                R = mul(R, M)
            M = mul(M, M)
            e >>= 1
        return R

    I = [1 << i for i in range(32)]
    N = 2 ** 32 - 1
    assert power(cols, N) == I
    for p in (3, 5, 17, 257, 65537):
        assert power(cols, N // p) != I


# --- CDF (scan) ---------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 1023, 1024, 1025, 4097, 70001])
def test_cdf_integer_scan_exact(n):
    """q_j = floor(w_j 2^(62-ceil(log2 n)) / max w) and cdf = cumsum(q) in
    uint64, bit for bit (numpy computes the same closed form independently)."""
    rng = np.random.default_rng(n)
    w = rng.lognormal(0, 3, size=n) ** 2
    w[rng.integers(0, n, size=max(1, n // 10))] = 0.0
    if n > 1:
        w[0] = 0.0
    cdf, tot, lp = O.cdf(w)
    L = int(np.ceil(np.log2(n))) if n > 1 else 0
    scale = np.ldexp(1.0, 62 - L) / w.max() if w.max() > 0 else 0.0
    q = np.floor(w * scale).astype(np.uint64)
    assert np.array_equal(cdf, np.cumsum(q, dtype=np.uint64))
    assert tot == int(q.sum()) and tot <= 2 ** 62
    nz = np.nonzero(q > 0)[0]
    assert lp == (nz[-1] if nz.size else 0)
    assert np.all(np.diff(cdf.astype(np.float64)) >= 0)


def test_cdf_law_and_nonfinite():
    """p_j = q_j / total differs from w_j / sum(w) by at most (n+1) 2^(L-62)."""
    rng = np.random.default_rng(1)
    n = 65536
    w = rng.lognormal(0, 3, size=n) ** 2
    cdf, tot, _ = O.cdf(w)
    q = np.diff(cdf.astype(object), prepend=0).astype(np.float64)
    p = q / float(tot)
    assert np.max(np.abs(p - w / w.sum())) <= (n + 1) * 2.0 ** (16 - 62)
    assert O.cdf_scale(np.array([1.0, np.nan])) < 0
    assert O.cdf_scale(np.array([1.0, np.inf])) < 0
    assert O.cdf_scale(np.zeros(4)) == 0.0
    cdf0, tot0, lp0 = O.cdf(np.zeros(4))
    assert tot0 == 0 and lp0 == 0


def test_draw_target_exact():
    rng = np.random.default_rng(17)
    for _ in range(2000):
        a, b = (int(v) for v in rng.integers(0, 2 ** 32, size=2))
        tot = int(rng.integers(1, 2 ** 62))
        t = O.draw_target(a, b, tot)
        assert t == (((a << 32) | b) * tot) >> 64
        assert 0 <= t < tot
    assert O.draw_target(0xFFFFFFFF, 0xFFFFFFFF, 2 ** 62) == 2 ** 62 - 1


# --- inverse-CDF search (pick_index, device_memory.h:529-543) ----------------
def test_search_matches_brute_force():
    rng = np.random.default_rng(2)
    w = rng.integers(0, 4, size=300).astype(np.float64)
    w[0] = 0.0
    w[-3:] = 0.0
    cdf, tot, lp = O.cdf(w)
    q = np.diff(cdf.astype(object), prepend=0)
    ts = [int(v) for v in rng.integers(0, tot, 3000)] + [int(v) for v in np.unique(cdf) if v < tot]
    ts += [0, tot - 1]
    for t in ts:
        j = O.search(cdf, t, lp)
        brute = next((k for k in range(cdf.size) if int(cdf[k]) > t), lp)
        assert j == brute
        assert q[j] > 0  # zero-weight columns are never picked
    # total = 0 (all-zero weights) falls back to the last positive column
    assert O.search(cdf, tot, lp) == lp


def test_sample_law_and_layout():
    rng = np.random.default_rng(3)
    n = 50
    w = rng.uniform(0.1, 3.0, n) ** 2
    w[[4, 17]] = 0.0
    c = 20000  # three segments of 8192
    idx, wgt, segc = O.sample(w, c, seed=11, iteration=5, trial=2)
    cd, tot, _ = O.cdf(w)
    qj = np.diff(cd.astype(object), prepend=0).astype(np.float64)
    counts = np.zeros(n)
    for g, u in enumerate(segc):
        seg = slice(g * O.SAMPLE_SEG, g * O.SAMPLE_SEG + int(u))
        ids = idx[seg]
        assert np.all(np.diff(ids.astype(np.int64)) > 0)  # unique, ascending
        mult = wgt[seg].astype(np.float64) * c * qj[ids] / float(tot)
        assert np.allclose(mult, np.rint(mult), rtol=1e-5)
        counts[ids] += np.rint(mult)
        seg_len = min(O.SAMPLE_SEG, c - g * O.SAMPLE_SEG)
        assert np.rint(mult).sum() == seg_len
    assert counts[[4, 17]].sum() == 0
    p = w / w.sum()
    keep = p > 0
    chi = stats.chisquare(counts[keep], c * p[keep])
    assert chi.pvalue > 1e-4
    # determinism and iteration dependence
    idx2, wgt2, _ = O.sample(w, c, seed=11, iteration=5, trial=2)
    assert np.array_equal(idx, idx2) and np.array_equal(wgt, wgt2)
    idx3, _, _ = O.sample(w, c, seed=11, iteration=6, trial=2)
    assert not np.array_equal(idx, idx3)


# --- column statistics (get_norms, device_memory.h:34-59) ---------------------
def test_colstats_closed_form():
    A = lowrank_plus_noise(100, 37, seed=4) + 3.0  # non-zero means
    cols = to_cols(A, ld=104)
    alpha, nrm2, cnrm2 = O.colstats(cols, d=100)
    A64 = A.astype(np.float64)
    mu = A64.mean(axis=0)
    assert np.allclose(alpha, mu.astype(np.float32), rtol=1e-7, atol=0)
    assert rel(nrm2, (A64 ** 2).sum(0)) < 1e-13
    assert rel(cnrm2, ((A64 - mu) ** 2).sum(0)) < 1e-10


def test_colstats_bf16():
    A = lowrank_plus_noise(64, 20, seed=5)
    bits = bf16_bits(A.T.copy())  # (n, d)
    alpha, nrm2, cnrm2 = O.colstats(bits, d=64, bf16=True)
    Ab = bf16_to_f32(bits).astype(np.float64).T
    assert rel(nrm2, (Ab ** 2).sum(0)) < 1e-13
    assert rel(cnrm2, ((Ab - Ab.mean(0)) ** 2).sum(0)) < 1e-10


# --- one AMM product -----------------------------------------------------------
def _setup(d=40, n=90, seed=6, shift=0.5):
    A = lowrank_plus_noise(d, n, seed=seed) + shift
    cols = to_cols(A)
    alpha, nrm2, cnrm2 = O.colstats(cols)
    A64 = A.astype(np.float64)
    a2 = float((alpha.astype(np.float64) ** 2).sum())
    z = (A64 @ alpha.astype(np.float64)).astype(np.float32)
    v = (z.astype(np.float64) - a2).astype(np.float32)
    return A, cols, alpha, nrm2, cnrm2, z, v


@pytest.mark.parametrize("proposed", [False, True])
def test_sanity_mode_equals_exact(proposed):
    """c = n, w = 1, all columns (device_memory.h:500-551): the sampled
    product is exact; for proposed this is the rank-2 identity
    A A^T dx = A'A'^T dx + (1^T dx) v + (z^T dx) 1."""
    A, cols, alpha, nrm2, cnrm2, z, v = _setup()
    d, n = A.shape
    rng = np.random.default_rng(7)
    x = rng.standard_normal(d).astype(np.float32)
    xm = rng.standard_normal(d).astype(np.float32)
    ym = rng.standard_normal(d).astype(np.float32)
    y = O.amm_product(cols, d, proposed, n, None, None, None, True, x, xm, ym, alpha, v, z)
    A64 = A.astype(np.float64)
    ref = A64 @ (A64.T @ (x.astype(np.float64) - xm)) + ym
    assert rel(y, ref) < 1e-6


def test_proposed_needs_the_rank2_terms():
    """Dropping the s*v + q*1 part (or the mean shift) breaks the identity."""
    A, cols, alpha, nrm2, cnrm2, z, v = _setup()
    d, n = A.shape
    x = np.random.default_rng(8).standard_normal(d).astype(np.float32)
    good = O.amm_product(cols, d, True, n, None, None, None, True, x, None, None, alpha, v, z)
    bad = O.amm_product(cols, d, True, n, None, None, None, True, x, None, None, alpha,
                        np.zeros_like(v), np.zeros_like(z))
    A64 = A.astype(np.float64)
    ref = A64 @ (A64.T @ x)
    assert rel(good, ref) < 1e-6 and rel(bad, ref) > 1e-2


def test_exact_product():
    A, cols, *_ = _setup()
    d, n = A.shape
    x = np.random.default_rng(9).standard_normal(d).astype(np.float32)
    A64 = A.astype(np.float64)
    assert rel(O.exact_product(cols, d, x), A64 @ (A64.T @ x)) < 1e-6


@pytest.mark.parametrize("proposed", [False, True])
def test_estimator_unbiased_with_closed_form_variance(proposed):
    """E[y_hat] = A A^T x, and E||y_hat - A A^T x||^2 =
    (1/c) (sum_j ||b_j||^2 (b_j^T x)^2 / p_j - ||B B^T x||^2) for the
    sampled part B (A, or the centred A' for proposed)."""
    A, cols, alpha, nrm2, cnrm2, z, v = _setup(d=24, n=60, seed=10, shift=2.0)
    d, n = A.shape
    w = cnrm2 if proposed else nrm2
    c = 8
    x = np.random.default_rng(11).standard_normal(d).astype(np.float32)
    A64 = A.astype(np.float64)
    truth = A64 @ (A64.T @ x)
    B = A64 - (alpha.astype(np.float64)[None, :] if proposed else 0.0)
    p = w / w.sum()
    bx = B.T @ x
    var = (np.sum((B ** 2).sum(0) * bx ** 2 / p) - np.linalg.norm(B @ bx) ** 2) / c
    N = 3000
    ys = np.empty((N, d))
    for it in range(N):
        idx, wgt, segc = O.sample(w, c, seed=5, iteration=it)
        ys[it] = O.amm_product(cols, d, proposed, c, idx, wgt, segc, False, x, None, None, alpha, v, z)
    mean = ys.mean(0)
    mc_var = np.mean(np.sum((ys - truth) ** 2, axis=1))
    assert np.linalg.norm(mean - truth) < 5 * np.sqrt(var / N)
    assert abs(mc_var / var - 1) < 0.15


# --- schedules (powerMethod_by_*.cuh) ------------------------------------------
def test_schedules():
    ex = lambda k, i, P=0, thr=0.0, cs=0.0, le=0: O.is_exact_step(k, i, P, thr, cs, le)  # noqa: E731
    # previous, -vr 5 (powerMethod_by_previous.cuh:135-138)
    assert [i for i in range(15) if ex("previousAMM", i, 5)] == [4, 9, 14]
    assert not any(ex("previousAMM", i) for i in range(10))  # no VR
    # adaptive: exact when cos of the previous iteration > 1 - thr
    assert ex("previousAMM", 3, 0, 0.1, 0.95) and not ex("previousAMM", 3, 0, 0.1, 0.85)
    # proposed gate i > last_exact + 1 (powerMethod_by_proposed.cuh:159)
    assert not ex("proposedAMM", 1, 2, le=0) and ex("proposedAMM", 3, 2, le=0)
    assert not ex("proposedAMM", 3, 2, le=3) and ex("proposedAMM", 5, 2, le=3)
    # proposed_v2 (powerMethod_by_proposed_v2.cuh:234,257,305): vr 4, step 12
    assert [i for i in range(12) if ex("proposedAMM_v2", i, 4)] == [3, 7, 11]
    assert all(ex("exact", i) for i in range(5))


def test_power_method_modes_follow_schedule():
    A = lowrank_plus_noise(32, 64, seed=12)
    cols = to_cols(A)
    r = O.power_method(cols, 32, "proposedAMM_v2", c=16, step=12, vr_period=4, seed=1)
    # 1 exact+memo, 2 previous, 3 proposed (amm_oracle.c mode_out)
    assert list(r["modes"]) == [2, 2, 2, 1, 3, 3, 3, 1, 3, 3, 3, 1]


# --- the power method against LAPACK -------------------------------------------
def test_exact_power_method_vs_eigvalsh():
    A = lowrank_plus_noise(96, 160, seed=13)
    lam = O.lambda_max_eigvalsh(A)
    r = O.power_method(to_cols(A), 96, "exact", step=80, seed=3)
    assert abs(r["eig"][-1] - lam) / lam < 1e-6
    assert abs(O.lambda_max_svd(A) - lam) / lam < 1e-12


def test_lanczos_oracle():
    A = heavy_columns(80, 300, seed=14).astype(np.float64)
    lam = O.lambda_max_eigvalsh(A)
    got = O.lanczos_lambda_max(lambda v: A @ (A.T @ v), 80, k=60)
    assert abs(got - lam) / lam < 1e-12


@pytest.mark.parametrize("kernel", ["previousAMM", "proposedAMM"])
def test_amm_power_method_converges_with_vr(kernel):
    """With periodic variance reduction the AMM iteration reaches the FP64
    eigenvalue (the control variate shrinks the sampling noise with dx)."""
    A = lowrank_plus_noise(64, 256, seed=15)
    lam = O.lambda_max_eigvalsh(A)
    r = O.power_method(to_cols(A), 64, kernel, c=64, step=60, vr_period=5, seed=4)
    assert abs(r["eig"][-1] - lam) / lam < 1e-3


def test_reference_style_sampler_unbiased():
    """The reference's own sampler restatement (xorshift slots, FP32 CDF)
    gives Rayleigh quotients around lambda_max."""
    A = lowrank_plus_noise(48, 200, seed=16, decay=2.0)
    lam = O.lambda_max_eigvalsh(A)
    x0 = O.x0(1, 0, 48)
    eig = O.ref_power_method(to_cols(A), 48, "previousAMM", 48, 200, 7, x0)
    assert abs(np.mean(eig[50:]) - lam) / lam < 0.05


def _lib_ref():
    import ctypes as C
    return C, O.lib()


def _xorshift_py(y: int) -> int:
    """Marsaglia's xorshift32 with shifts (13, 17, 5) in Python integers."""
    y ^= (y << 13) & 0xFFFFFFFF
    y ^= y >> 17
    y ^= (y << 5) & 0xFFFFFFFF
    return y


def test_reference_uniform23_bit_mapping():
    """orc_ref_uniform23 (get_rand, device_memory.h:506-517): the float
    (y & 0x7FFFFF | 0x3F800000) - 1 is exactly k / 2^23 with k the low 23
    bits of the xorshift output, and the state advances to that output.
    Checked against integer arithmetic, including outputs whose low bits are
    all zero (u = 0) and all one (u = 1 - 2^-23)."""
    C, L = _lib_ref()
    rng = np.random.default_rng(3)
    states = [int(s) for s in rng.integers(1, 2**32, 2000)]
    # states whose next output has low 23 bits all zero (u = 0) or all one
    # (u = 1 - 2^-23): invert the three xorshift steps (each is a bijection)
    def inv_shl(y, k):
        x = y
        for _ in range(32 // k + 1):
            x = y ^ ((x << k) & 0xFFFFFFFF)
        return x

    def inv_shr(y, k):
        x = y
        for _ in range(32 // k + 1):
            x = y ^ (x >> k)
        return x

    found = {}
    for y in (0x12800000, 0xFF800000, 0x007FFFFF, 0xA57FFFFF):
        s0 = inv_shl(inv_shr(inv_shl(y, 5), 17), 13)
        assert _xorshift_py(s0) == y
        found[y] = s0
    states += list(found.values())
    for s0 in states:
        st = np.array([s0], np.uint32)
        u = float(L.orc_ref_uniform23(C.c_void_p(st.ctypes.data)))
        y = _xorshift_py(s0)
        assert int(st[0]) == y
        assert u == (y & 0x7FFFFF) / 2.0**23
        assert 0.0 <= u < 1.0
        if s0 in found.values():
            assert u in (0.0, 1.0 - 2.0**-23)


def test_reference_pick_brute_force():
    """orc_ref_pick (pick_index, device_memory.h:519-545) against a linear
    scan: min{j : cdf_j > float32(cdf_{n-1} * u)}, n - 1 when no entry
    exceeds the target.  CDFs with leading and inner zero weights (ties),
    targets exactly on a CDF value, u = 0 and u just below 1."""
    C, L = _lib_ref()
    rng = np.random.default_rng(4)
    cases = [np.array([0, 0, 1, 1, 1, 3, 3, 6], np.float32),
             np.array([5], np.float32),
             np.array([0, 2], np.float32),
             np.cumsum(rng.choice([0.0, 0.25, 1.0, 3.5], size=37)).astype(np.float32)]
    for _ in range(20):
        w = rng.exponential(size=int(rng.integers(1, 300))).astype(np.float32)
        w[rng.random(w.size) < 0.3] = 0.0
        if w[-1] == 0.0:
            w[-1] = 1.0
        cases.append(np.cumsum(w, dtype=np.float32))
    for cdf in cases:
        n = cdf.size
        total = cdf[-1]
        us = [0.0, 1.0 - 2.0**-23] + [float(v) for v in rng.random(50)]
        # u that puts the target exactly on each distinct CDF value (when representable)
        for v in np.unique(cdf):
            u = np.float32(v / total) if total > 0 else np.float32(0)
            if np.float32(total * u) == v and u < 1:
                us.append(float(u))
        for u in us:
            u32 = np.float32(u)
            t = np.float32(total * u32)
            above = np.nonzero(cdf > t)[0]
            want = int(above[0]) if above.size else n - 1
            got = int(L.orc_ref_pick(C.c_void_p(cdf.ctypes.data), n, C.c_float(u32)))
            assert got == want, (cdf, float(u32), float(t), got, want)
