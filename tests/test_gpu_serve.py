"""The serving loop (vsmcx_serve_*) on the GPU, through the C ABI.

Jobs of the CV model and of the 10-D linear-Gaussian model are submitted while
others run, their observations arrive one step at a time, and every poll
launches one step of every job with an observation waiting (continuous
batching).  Checked here:
  * each job's results are the same alone and in the mix (bit-identical);
  * admission respects the HBM budget: with room for two jobs at a time, six
    jobs run in waves, each still giving its alone result;
  * a CV job served equals the same filter run by vsmcx_filter, bit for bit
    (and, in FP64, the CPU oracle within 1e-9);
  * LG jobs: the log-evidence estimate against the exact Kalman value
    (oracle/kalman.py), over 8 seeds, within 4 standard errors plus 0.05.
"""
import numpy as np
import pytest

import vsmc_amd as V
from kalman import kalman_logz

pytestmark = pytest.mark.gpu

Q, R, S0 = 0.1, 0.2, 0.25


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def A():
    return V.lg_matrix(10, seed=1)


def _serve(ctx, A, budget=8 << 30, max_jobs=64, precision=V.MIXED):
    return V.Serve(ctx, budget, max_jobs, 400, A, lg_q=Q, lg_r=R, lg_s0=S0, precision=precision)


def _jobs(A):
    """(model, n, T, seed, rseed, observations)"""
    out = []
    for k in range(6):
        model = V.MODEL_LG if k % 2 else V.MODEL_CV
        n = [1500, 40000, 3000, 2048, 100000, 700][k]
        T = [30, 12, 25, 40, 9, 33][k]
        y = V.lg_simulate(A, T, 50 + k, Q, R, S0) if model == V.MODEL_LG else V.cv_simulate(T, 50 + k)
        out.append((model, n, T, 300 + k, 900 + k, y))
    return out


def _run_alone(ctx, A, job, precision=V.MIXED):
    s = _serve(ctx, A, precision=precision)
    model, n, T, seed, rseed, y = job
    j = s.submit(model, n, T, seed, rseed)
    for t in range(T):
        s.observe(j, y[t])
        s.poll(wait=True)
    while s.job(j)[0] != 2:
        s.poll(wait=True)
    h = s.history(j, T)
    s.close()
    return h


def _run_mix(s, jobs, stagger):
    """Jobs arrive `stagger` polls apart; every poll each started job receives
    its next observation (if any)."""
    ids = [None] * len(jobs)
    fed = [0] * len(jobs)
    tick = 0
    while True:
        for k, job in enumerate(jobs):
            if ids[k] is None and tick >= stagger * k:
                model, n, T, seed, rseed, _ = job
                ids[k] = s.submit(model, n, T, seed, rseed)
            if ids[k] is not None and fed[k] < job[2]:
                s.observe(ids[k], job[5][fed[k]])
                fed[k] += 1
        s.poll(wait=True)
        tick += 1
        if all(i is not None and s.job(i)[0] == 2 for i in ids):
            break
        assert tick < 5000
    return [s.history(i, job[2]) for i, job in zip(ids, jobs)]


def test_serve_same_alone_and_in_mix(ctx, A):
    jobs = _jobs(A)
    say("serve: alone")
    alone = [_run_alone(ctx, A, j) for j in jobs]
    say("serve: mix")
    s = _serve(ctx, A)
    mix = _run_mix(s, jobs, stagger=3)
    st = s.stats()
    s.close()
    for k in range(len(jobs)):
        assert (alone[k] == mix[k]).all(), k
    assert st["jobs_done"] == len(jobs)
    assert st["steps"] == sum(j[2] for j in jobs)
    assert st["p99_us"] >= st["p50_us"] > 0


def test_serve_admission_under_budget(ctx, A):
    """Room for about two of the jobs at a time: they run in waves, results unchanged."""
    jobs = [(V.MODEL_CV, 60000, 10, 77 + k, 88 + k, V.cv_simulate(10, 80 + k)) for k in range(6)]
    alone = [_run_alone(ctx, A, j) for j in jobs]
    per_particle = 80 + 2 * 4 * 7  # arena bytes per particle (mixed, CV/LG average state)
    budget = int(2.2 * 61440 * per_particle)
    s = _serve(ctx, A, budget=budget)
    mix = _run_mix(s, jobs, stagger=0)
    st = s.stats()
    s.close()
    assert st["peak_particles"] <= st["capacity_particles"]
    assert st["peak_particles"] <= 2 * 61440
    for k in range(len(jobs)):
        assert (alone[k] == mix[k]).all(), k


@pytest.mark.parametrize("precision", [V.MIXED, V.FP64_FUSED])
def test_serve_cv_equals_single_filter(ctx, A, oracle, precision):
    """A served CV job runs the single-filter pipeline's tile and job bodies on
    the same tiles: its history equals vsmcx_filter's bit for bit (and, in
    FP64, the CPU oracle within 1e-9)."""
    n, T, seed, rseed = 5000, 20, 11, 12
    y = V.cv_simulate(T, 3)
    h = _run_alone(ctx, A, (V.MODEL_CV, n, T, seed, rseed, y), precision=precision)
    f = V.Filter(ctx, V.cv_desc(n, T, precision=precision, scheme=V.SYSTEMATIC, seed=seed,
                                resample_seed=rseed))
    f.set_obs(y)
    f.run(T)
    f.finish()
    hf = f.history()
    f.close()
    assert (h == hf).all()
    if precision == V.FP64_FUSED:
        _, _, rec = oracle.cv_run(n, y, seed, rseed, V.SYSTEMATIC, 0.5)
        assert (h["resampled"] == rec["resampled"]).all()
        np.testing.assert_allclose(h["log_z"], rec["log_z"], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision", [V.FP64_FUSED, V.MIXED])
def test_serve_lg_logz_vs_kalman(ctx, A, precision):
    T, n, R_ = 40, 1 << 14, 8
    y = V.lg_simulate(A, T, 123, Q, R, S0)
    exact = kalman_logz(A, y, Q, R, S0)[-1]
    s = _serve(ctx, A, precision=precision)
    jobs = [(V.MODEL_LG, n, T, 7000 + r, 8000 + r, y) for r in range(R_)]
    hs = _run_mix(s, jobs, stagger=1)
    s.close()
    lz = np.array([h["log_z"][-1] for h in hs])
    se = lz.std(ddof=1) / np.sqrt(R_)
    say("LG logZ", precision, lz.mean(), exact, se)
    assert abs(lz.mean() - exact) < 4 * se + 0.05


def test_serve_lg_one_dimensional(ctx):
    """A one-dimensional linear-Gaussian state: the monitor's second coordinate
    is 0 at every step (the state has none, as the move's own monitor treats
    it), and log Z matches the Kalman filter's within 4 standard errors plus
    0.05 over 8 seeds."""
    A1 = V.lg_matrix(1, seed=3)
    T, n, R_ = 30, 1 << 12, 8
    y = V.lg_simulate(A1, T, 77, Q, R, S0)
    exact = kalman_logz(A1, y, Q, R, S0)[-1]
    s = _serve(ctx, A1)
    hs = _run_mix(s, [(V.MODEL_LG, n, T, 100 + r, 200 + r, y) for r in range(R_)], stagger=1)
    s.close()
    for h in hs:
        assert np.all(h["mean_y"] == 0.0)
        assert np.all(np.isfinite(h["mean_x"]))
    lz = np.array([h["log_z"][-1] for h in hs])
    se = lz.std(ddof=1) / np.sqrt(R_)
    say("LG 1-D logZ", lz.mean(), exact, se)
    assert abs(lz.mean() - exact) < 4 * se + 0.05
