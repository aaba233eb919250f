"""The many-matrix configuration at full size: 256 matrices with
d ~ U[512, 8192] and n ~ U[1024, 32768] (the 1024 x 2048 construction:
s_i = 10/i, Haar factors, 0.01 N(0,1) noise, truncated to rank 64), stored
in BF16, 16 sampling seeds each stacked as one right-hand-side block,
previousAMM with c = n/8 and an exact step every 3 iterations, 150
iterations, all in one batched run.

* every (matrix, trial) ends within 1e-5 of the FP64 Lanczos value of its
  stored BF16 matrix and reaches 1e-3;
* for 8 matrices the Lanczos value itself agrees with FP64 eigvalsh of the
  Gram matrix to 1e-10;
* the trials of 3 matrices follow the single-matrix run with the same seed
  and trial word (same samples): identical step kinds, Rayleigh quotients
  within 1e-4.
The reference solves one matrix and one right-hand side per run
(AMM.cu:7-72); the loop each trial follows is powerMethod_by_previous.cuh:
100-210.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

COUNT, TRIALS, ITERS, VR = 256, 16, 150, 3


def banner(msg):
    print(f"\n[gpu-case] {msg}", flush=True)


def relmax(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.fixture(scope="module")
def c4():
    from eigenvaluecomputation_by_amm_amd.workloads import (lambda_max_lanczos, lowrank_plus_noise_fast,
                                                            uneven_problem_shapes)
    shapes = uneven_problem_shapes(COUNT)
    banner(f"C4 generate {COUNT} matrices ({sum(d * n for d, n in shapes) / 1e9:.1f} G elements) + Lanczos")
    mats, lams, changes = [], [], []
    for i, (d, n) in enumerate(shapes):
        M = lowrank_plus_noise_fast(d, n, seed=1000 + i).to(torch.bfloat16)
        lam, ch = lambda_max_lanczos(M, d, k=40)
        mats.append(M)
        lams.append(lam)
        changes.append(ch)
        if i % 64 == 63:
            print(f"[gpu-case] C4 generated {i + 1}", flush=True)
    yield dict(shapes=shapes, mats=mats, lams=np.array(lams), changes=np.array(changes))
    del mats
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_c4_every_solve_against_lanczos(c4):
    from eigenvaluecomputation_by_amm_amd import AmmBatch, AmmContext, RunParams
    banner("C4 batched previousAMM c=n/8 vr=3 x16, 150 iterations")
    shapes, mats, lams = c4["shapes"], c4["mats"], c4["lams"]
    assert c4["changes"].max() < 1e-12
    probs = [dict(A=mats[i], c=max(1, n // 8), seed=77 + i) for i, (d, n) in enumerate(shapes)]
    params = RunParams(kernel="previousAMM", vr_period=VR, step=ITERS, trial=0)
    with AmmBatch(probs, trials=TRIALS) as b:
        r = b.power_method(params)
        modes = b.last_modes(ITERS)
    err = np.abs(r.eig - lams[:, None, None]) / lams[:, None, None]  # [problem][trial][step]
    assert err[:, :, -1].max() < 1e-5, err[:, :, -1].max()
    assert (err <= 1e-3).any(axis=2).all()
    banner("C4 trials of 3 matrices against the single-matrix path")
    for i in (0, 101, 202):
        d, n = shapes[i]
        with AmmContext(mats[i], d=d) as ctx:
            for k in (0, 7, 15):
                s = ctx.power_method(RunParams(kernel="previousAMM", c=max(1, n // 8), vr_period=VR,
                                               step=ITERS, seed=77 + i, trial=k))
                assert np.array_equal(ctx.last_modes(ITERS), modes[i, k])
                assert relmax(r.eig[i, k], s.eig) < 1e-4, (i, k, relmax(r.eig[i, k], s.eig))


@pytest.mark.timeout(900)
def test_c4_lanczos_against_eigvalsh(c4):
    """The ground truth of the batch: FP64 Lanczos against FP64 eigvalsh of
    A A^T for 8 matrices spread over the row counts up to 4096."""
    banner("C4 Lanczos vs eigvalsh, 8 matrices")
    shapes, mats, lams = c4["shapes"], c4["mats"], c4["lams"]
    small = sorted((i for i, (d, _) in enumerate(shapes) if d <= 4096), key=lambda i: shapes[i][0])
    pick = [small[int(round(k * (len(small) - 1) / 7))] for k in range(8)]
    for i in pick:
        d, n = shapes[i]
        M64 = mats[i][:, :d].double()
        lam = float(torch.linalg.eigvalsh(M64.T @ M64)[-1])
        del M64
        assert abs(lams[i] - lam) / lam < 1e-10, (i, d, n, lams[i], lam)
