"""Each workload configuration run at its stated size on the GPU, through the
C ABI, and checked against the oracle or against the single-filter path:

  C0  the pf filter, N = 10^4, T = 100, FP64, Philox4x32-10, systematic at
      ESS < N/2: against the FP64 oracle (example/pf restated);
  C2  the GMM tempered sampler, K = 4, 1000 points, 5 MH sweeps:
      FP64 on the GPU against the oracle at N = 4096, and at N = 2^20 the
      mixed-precision log evidence against the FP64 GPU run's;
  C3  4096 batched filters, N log-uniform in [2^8, 2^16], T uniform in
      [50, 500]: 16 sampled filters bit-identical to vsmcx_filter's run
      (mixed), 4 sampled FP64 filters against the oracle;
  C4  the serving loop: Poisson arrivals at 2000/s until 500 jobs, CV and
      10-D linear-Gaussian, N log-uniform in [2^10, 2^20], T uniform in
      [20, 1000], 200 GB budget: 8 sampled jobs equal to their runs alone.
(C1, N = 2^24, is covered in test_gpu_core.py.)
"""
import time

import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu


def say(*a):
    print("[configs]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def test_c0_fp64_vs_oracle(ctx, oracle):
    N, T = 10000, 100
    obs = V.cv_simulate(T, 2024)
    say("C0", N, T)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.FP64_FUSED, scheme=V.SYSTEMATIC, seed=42,
                                resample_seed=4242))
    try:
        f.set_obs(obs)
        f.run(T)
        f.finish()
        h = f.history()
        st, lw = f.state()
    finally:
        f.close()
    ost, olw, rec = oracle.cv_run(N, obs, 42, 4242, V.SYSTEMATIC, 0.5)
    assert (h["resampled"] == rec["resampled"]).all()
    assert (h["draws"] == rec["draws"]).all()
    for k in ("ess", "log_z", "mean_x", "mean_y"):
        np.testing.assert_allclose(h[k], rec[k], rtol=1e-9, atol=1e-9, err_msg=k)
    np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)


def _gmm(ctx, y, n, precision, seed):
    g = V.Gmm(ctx, y, n, components=4, sweeps=5, precision=precision, seed=seed,
              resample_seed=seed + 1)
    try:
        g.initialize()
        g.iterate()
        h = g.history()
        st = g.state() if n <= 8192 else None
    finally:
        g.close()
    return h, st


def test_c2_fp64_vs_oracle_1000_points_5_sweeps(ctx, oracle):
    y = V.gmm_simulate(1000, 23)
    n = 4096
    say("C2 fp64 vs oracle", n)
    h, (p, llh, lw) = _gmm(ctx, y, n, V.FP64_FUSED, 24)
    op, ollh, olw, rec = oracle.gmm_run(n, y, K=4, sweeps=5, seed=24, resample_seed=25,
                                        threads=16)
    assert len(h) == len(rec) and h["alpha"][-1] == 1.0
    for k in ("resampled", "bisections", "accept"):
        assert (h[k] == rec[k]).all(), k
    for k in ("alpha", "alpha_inc", "ess", "log_z", "path_integrand", "path_log_z"):
        np.testing.assert_allclose(h[k], rec[k], rtol=1e-9, atol=1e-12, err_msg=k)
    np.testing.assert_allclose(p, op, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(llh, ollh, rtol=1e-9)


def test_c2_full_size_mixed_vs_fp64(ctx):
    """2^20 particles x 1000 points x 5 sweeps.  The mixed and the FP64 runs
    share seeds but are different Monte Carlo realisations once an MH
    decision or an ancestor differs, so their log evidences agree to the
    estimator's spread: the bound is 4 standard errors of the difference,
    with the standard error of one run estimated from 4 mixed runs on other
    seeds, plus 0.02 for the FP32 likelihood's bias (1e-5 relative per
    point, summed over the tempering path, is below 0.02 here)."""
    y = V.gmm_simulate(1000, 23)
    n = 1 << 20
    say("C2 full size fp64")
    h64, _ = _gmm(ctx, y, n, V.FP64_FUSED, 24)
    zs = []
    for s in range(5):
        say("C2 full size mixed", s)
        h32, _ = _gmm(ctx, y, n, V.MIXED, 24 + 10 * s)
        assert h32["alpha"][-1] == 1.0
        zs.append(h32["log_z"][-1])
    se = float(np.std(zs[1:], ddof=1))
    d = zs[0] - h64["log_z"][-1]
    say("C2 log Z mixed", zs, "fp64", h64["log_z"][-1], "sd", se)
    assert abs(d) < 4 * np.sqrt(2) * se + 0.02, (d, se)
    assert 20 <= len(h64) <= 80


def _c3_workload(nfilters, seed):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(2 ** 8), np.log(2 ** 16), nfilters)).astype(np.uint64)
    Ts = rng.integers(50, 501, nfilters).astype(np.uint64)
    obs = V.cv_simulate_many(Ts, seed + 1)
    seeds = 100000 + np.arange(nfilters, dtype=np.uint64)
    rseeds = 200000 + np.arange(nfilters, dtype=np.uint64)
    return ns, Ts, obs, seeds, rseeds


def test_c3_4096_filters_equal_single_filters(ctx):
    ns, Ts, obs, seeds, rseeds = _c3_workload(4096, 13)
    say("C3 mixed batch", int(ns.sum()))
    b = V.Batch(ctx, ns, Ts, seeds, rseeds, precision=V.MIXED, scheme=V.SYSTEMATIC)
    try:
        for f in range(ns.size):
            b.set_obs(f, obs[f])
        b.tune(1)
        b.run()
        pick = np.random.default_rng(5).choice(ns.size, 16, replace=False)
        for f in pick:
            f = int(f)
            hb = b.history(f)
            sb, lwb = b.state(f)
            fl = V.Filter(ctx, V.cv_desc(int(ns[f]), int(Ts[f]), precision=V.MIXED,
                                         scheme=V.SYSTEMATIC, seed=int(seeds[f]),
                                         resample_seed=int(rseeds[f])))
            try:
                fl.set_obs(obs[f])
                fl.run(int(Ts[f]))
                fl.finish()
                assert (fl.history() == hb).all(), f
                st, lw = fl.state()
            finally:
                fl.close()
            assert (st == sb).all() and (lw == lwb).all(), f
    finally:
        b.close()


def test_c3_4096_filters_fp64_vs_oracle(ctx, oracle):
    ns, Ts, obs, seeds, rseeds = _c3_workload(4096, 13)
    say("C3 fp64 batch")
    b = V.Batch(ctx, ns, Ts, seeds, rseeds, precision=V.FP64_FUSED, scheme=V.SYSTEMATIC)
    try:
        for f in range(ns.size):
            b.set_obs(f, obs[f])
        b.run()
        # four filters spread over the sizes, the oracle's runs kept to seconds
        order = np.argsort(ns * Ts)
        for f in order[[0, 1000, 2500, 3500]]:
            f = int(f)
            h = b.history(f)
            st, lw = b.state(f)
            ost, olw, rec = oracle.cv_run(int(ns[f]), obs[f], int(seeds[f]), int(rseeds[f]),
                                          V.SYSTEMATIC, 0.5)
            assert (h["resampled"] == rec["resampled"]).all(), f
            assert (h["draws"] == rec["draws"]).all(), f
            for k in ("ess", "log_z", "mean_x", "mean_y"):
                np.testing.assert_allclose(h[k], rec[k], rtol=1e-9, atol=1e-9, err_msg=k)
            np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
    finally:
        b.close()


def test_c4_500_poisson_jobs_equal_alone(ctx):
    rate, njobs, seed = 2000.0, 500, 17
    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, njobs))
    models = rng.integers(0, 2, njobs)
    ns = np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), njobs)).astype(np.uint64)
    Ts = rng.integers(20, 1001, njobs)
    A = V.lg_matrix(10, seed=seed)
    y_cv = V.cv_simulate_many(Ts, seed + 1)
    y_lg = V.lg_simulate_many(A, Ts, seed + 2)
    data = [y_lg[k] if models[k] == V.MODEL_LG else y_cv[k] for k in range(njobs)]
    say("C4 mix", njobs, "jobs")
    s = V.Serve(ctx, 200e9, 8192, 1000, A, precision=V.MIXED)
    ids, fed = {}, {}
    try:
        t0 = time.perf_counter()
        nxt = 0
        done = 0
        while done < njobs:
            now = time.perf_counter() - t0
            while nxt < njobs and t_arr[nxt] <= now:
                j = s.submit(int(models[nxt]), int(ns[nxt]), int(Ts[nxt]), 500000 + nxt,
                             600000 + nxt)
                ids[nxt] = j
                fed[j] = 0
                nxt += 1
            ready = s.ready()
            if ready.size:
                kk = {j: k for k, j in ids.items()}
                ys = [data[kk[int(j)]][fed[int(j)]] for j in ready]
                s.observe_many(ready, ys)
                for j in ready:
                    fed[int(j)] += 1
            launched = s.poll(wait=True)
            done = sum(1 for k, j in ids.items() if fed[j] == Ts[k] and s.job(j)[0] == 2)
            if launched == 0 and nxt < njobs:
                time.sleep(max(0.0, min(1e-3, t_arr[nxt] - (time.perf_counter() - t0))))
        st = s.stats()
        assert st["jobs_done"] == njobs
        assert st["peak_particles"] <= st["capacity_particles"]
        pick = [int(k) for k in np.random.default_rng(3).choice(njobs, 8, replace=False)]
        mixed = {k: s.history(ids[k], int(Ts[k])) for k in pick}
    finally:
        s.close()
    for k in pick:
        say("C4 alone", k, int(models[k]), int(ns[k]), int(Ts[k]))
        a = V.Serve(ctx, 16 << 30, 4, 1000, A, precision=V.MIXED)
        try:
            j = a.submit(int(models[k]), int(ns[k]), int(Ts[k]), 500000 + k, 600000 + k)
            for t in range(int(Ts[k])):
                a.observe(j, data[k][t])
                a.poll(wait=True)
            while a.job(j)[0] != 2:
                a.poll(wait=True)
            alone = a.history(j, int(Ts[k]))
        finally:
            a.close()
        assert (alone == mixed[k]).all(), k
