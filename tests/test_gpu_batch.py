"""Batched filters (vsmcx_batch_*) on the GPU, through the C ABI.

Each filter of a batch is the pf model of vsmcx_filter with its own particle
count, horizon, observations and seeds; the batch advances them all with one
persistent launch per run.  Checked here:
  * FP64 batch vs the CPU oracle, per filter: resampling decisions and engine
    draws identical, ESS / log-evidence / means / final state within 1e-9
    relative (the oracle restates example/pf and the resampling of vSMC);
  * a filter's results do not depend on the launch shape, on the filter order
    or on the other filters in the batch: bit-identical histories;
  * mixed precision: first step (before any resample) coupled to FP64 within
    1e-5, as for the single-filter pipeline;
  * runs advance as far as the observations allow (serving-style feeding).
"""
import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _obs(T, seed):
    return V.cv_simulate(T, seed)


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def _make(ctx, ns, Ts, scheme, precision, base=0):
    k = len(ns)
    seeds = [1000 + base + 7 * i for i in range(k)]
    rseeds = [5000 + base + 11 * i for i in range(k)]
    b = V.Batch(ctx, ns, Ts, seeds, rseeds, precision=precision, scheme=scheme)
    return b, seeds, rseeds


@pytest.mark.parametrize("scheme", [V.SYSTEMATIC, V.STRATIFIED])
def test_batch_fp64_vs_oracle(ctx, oracle, scheme):
    ns = [1, 5, 300, 2048, 2049, 7000, 20000]
    Ts = [3, 9, 25, 17, 12, 20, 15]
    b, seeds, rseeds = _make(ctx, ns, Ts, scheme, V.FP64_FUSED)
    obs = [_obs(T, 40 + i) for i, T in enumerate(Ts)]
    for f in range(len(ns)):
        b.set_obs(f, obs[f])
    say("batch fp64", scheme)
    b.run()
    for f, (n, T) in enumerate(zip(ns, Ts)):
        h = b.history(f)
        assert len(h) == T
        st, lw = b.state(f)
        ost, olw, rec = oracle.cv_run(n, obs[f], seeds[f], rseeds[f], scheme, 0.5)
        assert (h["resampled"] == rec["resampled"]).all(), f
        assert (h["draws"] == rec["draws"]).all(), f
        np.testing.assert_allclose(h["ess"], rec["ess"], rtol=1e-9)
        np.testing.assert_allclose(h["log_z"], rec["log_z"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(h["mean_x"], rec["mean_x"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(h["mean_y"], rec["mean_y"], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)
    b.close()


@pytest.mark.parametrize("precision", [V.MIXED, V.FP64_FUSED])
def test_batch_independent_of_launch_and_company(ctx, precision):
    """The same filter alone, in a crowd, in another order and under other
    launch shapes gives bit-identical histories and states."""
    rng = np.random.default_rng(3)
    k = 64
    ns = np.exp(rng.uniform(np.log(256), np.log(40000), k)).astype(np.uint64)
    Ts = rng.integers(5, 30, k).astype(np.uint64)
    obs = [_obs(int(T), 100 + i) for i, T in enumerate(Ts)]
    seeds = 3000 + np.arange(k, dtype=np.uint64)
    rseeds = 9000 + np.arange(k, dtype=np.uint64)

    def run(idx, workgroups, order, chunks, strategy=V.BATCH_PERSISTENT):
        b = V.Batch(ctx, ns[idx], Ts[idx], seeds[idx], rseeds[idx], precision=precision)
        for j, f in enumerate(idx):
            b.set_obs(j, obs[f])
        b.set_strategy(strategy)
        b.set_launch(workgroups, order)
        for c in chunks:
            b.run(c)
        b.run()
        out = {int(f): (b.history(j), b.state(j)) for j, f in enumerate(idx)}
        b.close()
        return out

    say("batch crowd A")
    a = run(np.arange(k), 0, 0, [])
    say("batch crowd B (reversed, 7 workgroups, 3+5 steps first)")
    bb = run(np.arange(k)[::-1].copy(), 7, 1, [3, 5])
    say("batch alone")
    alone = run(np.array([17]), 1, 0, [1])
    say("batch crowd C (ticks, 4+2 steps first)")
    cc = run(np.arange(k), 0, 0, [4, 2], strategy=V.BATCH_TICKS)
    for f in range(k):
        ha, (sa, la) = a[f]
        for other in (bb, cc):
            hb, (sb, lb) = other[f]
            assert (ha == hb).all(), f
            assert (sa == sb).all() and (la == lb).all(), f
    h1, (s1, l1) = alone[17]
    assert (h1 == a[17][0]).all() and (s1 == a[17][1][0]).all() and (l1 == a[17][1][1]).all()


@pytest.mark.parametrize("precision", [V.MIXED, V.FP64_FUSED])
@pytest.mark.parametrize("strategy", [V.BATCH_PERSISTENT, V.BATCH_TICKS])
def test_batch_equals_single_filter(ctx, precision, strategy):
    """Every engine runs the same tile bodies and the same fixed-order
    reductions: a batched filter's history and final state are vsmcx_filter's
    for the same filter, bit for bit, under either strategy."""
    ns, Ts = [1500, 5000, 9000], [20, 15, 9]
    seeds, rseeds = [31, 32, 33], [41, 42, 43]
    obs = [_obs(T, 60 + i) for i, T in enumerate(Ts)]
    b = V.Batch(ctx, ns, Ts, seeds, rseeds, precision=precision, scheme=V.STRATIFIED)
    for f in range(3):
        b.set_obs(f, obs[f])
    b.set_strategy(strategy)
    b.run()
    for f in range(3):
        fl = V.Filter(ctx, V.cv_desc(ns[f], Ts[f], precision=precision, scheme=V.STRATIFIED,
                                     seed=seeds[f], resample_seed=rseeds[f]))
        fl.set_obs(obs[f])
        fl.run(Ts[f])
        fl.finish()
        hf = fl.history()
        st, lw = fl.state()
        fl.close()
        assert (b.history(f) == hf).all(), f
        bs, blw = b.state(f)
        assert (bs == st).all() and (blw == lw).all(), f
    b.close()


def test_batch_mixed_first_step_coupled(ctx, oracle):
    ns, Ts = [10000, 3000], [1, 1]
    b, seeds, rseeds = _make(ctx, ns, Ts, V.SYSTEMATIC, V.MIXED, base=5)
    obs = [_obs(1, 5), _obs(1, 6)]
    for f in range(2):
        b.set_obs(f, obs[f])
    b.run()
    for f in range(2):
        h = b.history(f)
        _, _, rec = oracle.cv_run(ns[f], obs[f], seeds[f], rseeds[f], V.SYSTEMATIC, 0.5)
        assert abs(h["ess"][0] - rec["ess"][0]) <= 1e-5 * rec["ess"][0]
        assert abs(h["log_z"][0] - rec["log_z"][0]) <= 1e-5
    b.close()


def test_batch_feeds_as_observations_arrive(ctx):
    """Runs stop where the observations stop and continue when more arrive; the
    result equals a run with every observation present from the start."""
    ns, Ts = [500, 4100], [12, 9]
    obs = [_obs(12, 1), _obs(9, 2)]
    b1, _, _ = _make(ctx, ns, Ts, V.SYSTEMATIC, V.FP64_FUSED)
    for f in range(2):
        b1.set_obs(f, obs[f])
    b1.run()
    b2, _, _ = _make(ctx, ns, Ts, V.SYSTEMATIC, V.FP64_FUSED)
    t = [0, 0]
    for step in range(12):
        for f in range(2):
            if t[f] < Ts[f]:
                b2.set_obs(f, obs[f][t[f]:t[f] + 1], first=t[f])
                t[f] += 1
        b2.run()
        for f in range(2):
            assert b2.steps(f) == t[f]
    for f in range(2):
        assert (b1.history(f) == b2.history(f)).all()
        s1, l1 = b1.state(f)
        s2, l2 = b2.state(f)
        assert (s1 == s2).all() and (l1 == l2).all()
    b1.close()
    b2.close()


def test_batch_tuner_keeps_results(ctx):
    ns = [3000] * 8 + [300] * 8
    Ts = [6] * 16
    b, _, _ = _make(ctx, ns, Ts, V.SYSTEMATIC, V.MIXED)
    ref, _, _ = _make(ctx, ns, Ts, V.SYSTEMATIC, V.MIXED)
    for f in range(16):
        o = _obs(6, 70 + f)
        b.set_obs(f, o)
        ref.set_obs(f, o)
    wg, order = b.tune(2)  # candidate strategies and shapes, each on the same two steps
    assert wg >= 1 and order in (0, 1) and b.strategy() in (V.BATCH_PERSISTENT, V.BATCH_TICKS)
    # tuning restores the state after every candidate: no filter advanced
    assert all(b.steps(f) == 0 for f in range(16))
    rep = b.tune_report()
    assert len(rep) >= 2 and all(ms > 0 for *_, ms in rep)
    assert {r[0] for r in rep} == {V.BATCH_PERSISTENT, V.BATCH_TICKS}
    best = min(rep, key=lambda r: r[3])
    assert b.strategy() == best[0] and (best[0] == V.BATCH_TICKS or wg == best[1])
    b.run()
    ref.run()
    for f in range(16):
        assert (b.history(f) == ref.history(f)).all()
    b.close()
    ref.close()


def test_batch_rejects_unsupported(ctx):
    with pytest.raises(V.VsmcxError):
        V.Batch(ctx, [100], [5], [1], [2], scheme=V.MULTINOMIAL)
    with pytest.raises(V.VsmcxError):
        V.Batch(ctx, [(1 << 20) + 1], [5], [1], [2])
