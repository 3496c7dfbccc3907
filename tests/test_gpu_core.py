"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bit-exact: RNG blocks, the exact left fold, replication counts, cfrp_trans
ancestor layout and resample-engine state for all six schemes.
Tolerance (stated per assertion): filter log-weights, ESS, log-evidence, means.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import vsmc_amd as V
from oracle_ctypes import cv_param

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


# ---------------------------------------------------------------- RNG
def test_philox_fill(ctx, oracle):
    say("philox fill")
    rng = np.random.default_rng(0)
    for _ in range(8):
        key = rng.integers(0, 2**32, 2, dtype=np.uint64).astype(np.uint32)
        ctr = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        ctr[0] = 2**32 - 3  # carry into word 1 inside the fill
        out = ctx.philox4x32_fill(key, ctr, 257)
        for k in (0, 1, 2, 3, 100, 256):
            c = ctr.astype(np.uint64).copy()
            v = int(c[0]) + k + 1
            c0, carry = v % 2**32, v // 2**32
            want = oracle.philox4x32_block([c0, int(c[1]) + carry, c[2], c[3]], key)
            assert (out[k] == want).all(), k


def test_philox_fill_matches_engine_stream(ctx, oracle):
    out = ctx.philox4x32_fill([1234, 0], [0, 0, 0, 0], 1000)
    assert (out.reshape(-1) == oracle.ph32_stream(1234, 4000)).all()


def test_threefry_fills(ctx, oracle):
    say("threefry fills")
    out = ctx.threefry4x64_fill([77, 0, 0, 0], [0, 0, 0, 0], 1000)
    assert (out.reshape(-1) == oracle.tf64_stream(77, 4000)).all()
    out = ctx.threefry4x32_fill([77, 0, 0, 0], [0, 0, 0, 0], 1000)
    assert (out.reshape(-1) == oracle.tf32_stream(77, 4000)).all()
    ctr = [2**64 - 2, 2**64 - 1, 3, 0]
    out = ctx.threefry4x64_fill([5, 6, 7, 8], ctr, 5)
    for k in range(5):
        v = ctr[0] + k + 1
        c = [v % 2**64, (ctr[1] + v // 2**64) % 2**64, ctr[2] + ((ctr[1] + v // 2**64) // 2**64),
             0]
        assert (out[k] == oracle.threefry4x64_block(c, [5, 6, 7, 8])).all()


def test_stream_probe(ctx):
    """The streaming probe reports rates between a floor no HBM part is below
    and the published 8 TB/s peak, and gives its buffers back."""
    say("stream probe")
    before = ctx.allocated()
    r = ctx.stream_probe(1 << 28, 3)
    say("stream probe", r)
    for k in ("copy_GBps", "read_GBps", "write_GBps"):
        assert 500.0 < r[k] < 8000.0, (k, r[k])
    assert ctx.allocated() == before
    with pytest.raises(V.VsmcxError):
        ctx.stream_probe(1 << 10, 3)


# ---------------------------------------------------------------- fold
def _weights(kind, n, rng):
    if kind == "flat":
        w = rng.random(n)
    elif kind == "heavy":
        w = np.exp(3 * rng.standard_normal(n))
    elif kind == "zeros":
        w = rng.random(n) * (rng.random(n) < 0.3)
        w[0] = 0.0
    elif kind == "degenerate":  # a few heavy particles, a sea of ~1e-16 weights
        w = np.full(n, 1e-16) * rng.random(n)
        w[rng.integers(0, max(1, n // 3), 5)] = 1.0
    elif kind == "ties":
        w = np.ldexp(1.0, -rng.integers(0, 60, n).astype(np.int64))
    elif kind == "range":
        w = np.exp(20 * rng.standard_normal(n))
    else:
        raise ValueError(kind)
    if w.sum() == 0:
        w[-1] = 1.0
    return w / w.sum() if kind != "ties" else w


@pytest.mark.parametrize("kind", ["flat", "heavy", "zeros", "degenerate", "ties", "range"])
def test_fold_prefix_exact(ctx, oracle, kind):
    rng = np.random.default_rng(hash(kind) % 2**32)
    for n in (1, 2, 3, 4095, 4096, 4097, 50000, 1 << 20):
        say("fold", kind, n)
        x = _weights(kind, n, rng)
        got = ctx.fold_prefix(x)
        want = oracle.fold_prefix(x)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (kind, n, bad[:5], got[bad[:5]], want[bad[:5]])


# ---------------------------------------------------------------- resample
SIZES = (1, 2, 3, 17, 4096, 4097, 30000)


def _engine_equal(eng, oeng):
    """Device engine state (ctr, key, index, buffer) equals the oracle engine's."""
    assert list(eng.ctr) == list(oeng.ctr)
    assert list(eng.key) == list(oeng.key)
    assert int(eng.index) == int(oeng.index)
    if eng.index < 4:
        assert list(eng.buffer) == list(oeng.buf)


@pytest.mark.parametrize("scheme", range(6))
def test_resample_compat_bitexact(ctx, oracle, scheme):
    rng = np.random.default_rng(100 + scheme)
    for n in SIZES:
        for kind in ("flat", "heavy", "zeros", "degenerate"):
            say("resample", V.SCHEMES and scheme, n, kind)
            w = _weights(kind, n, rng)
            seed = int(rng.integers(1, 2**31))
            eng = V.ThreefryState.seeded(seed)
            rep, cf = ctx.resample(scheme, w, eng)
            orep, draws, oeng = oracle.resample(scheme, w, seed)
            assert (rep == orep).all(), (scheme, n, kind)
            assert (cf == oracle.cfrp_trans(orep)).all(), (scheme, n, kind)
            _engine_equal(eng, oeng)


@pytest.mark.parametrize("scheme", range(6))
def test_resample_engine_continues(ctx, oracle, scheme):
    """Several resamples in a row on one engine (buffer index carried across)."""
    rng = np.random.default_rng(7 + scheme)
    eng = V.ThreefryState.seeded(99)
    oeng = oracle.tf64_engine(99)
    for r in range(6):
        n = int(rng.integers(1, 3000))
        w = _weights("heavy", n, rng)
        rep, cf = ctx.resample(scheme, w, eng)
        orep, _, oeng = oracle.resample(scheme, w, engine=oeng)
        assert (rep == orep).all(), (scheme, r, n)
        _engine_equal(eng, oeng)


@pytest.mark.parametrize("M,N", [(10, 3), (3, 10), (1000, 4100), (5000, 17)])
def test_resample_m_ne_n(ctx, oracle, M, N):
    rng = np.random.default_rng(M * 7 + N)
    for scheme in range(6):
        w = _weights("flat", M, rng)
        eng = V.ThreefryState.seeded(5)
        rep, cf = ctx.resample(scheme, w, eng, N=N)
        oe = oracle.tf64_engine(5)
        orep = np.zeros(M, np.uint64)
        oracle.lib.orc_resample(scheme, M, N, C.byref(oe), w, orep)
        assert (rep == orep).all(), (scheme, M, N)
        assert int(rep.sum()) == N
        assert (cf == oracle.cfrp_trans(orep, N)).all()
        _engine_equal(eng, oe)


# ---------------------------------------------------------------- filter
def _obs(T, seed=3):
    return V.cv_simulate(T, seed)


@pytest.mark.parametrize("scheme", range(6))
def test_filter_fp64_matches_oracle(ctx, oracle, scheme):
    """FP64 fused filter vs the FP64 oracle run: identical resampling decisions
    and engine draws; ESS, log-evidence and means to rtol 1e-9 (GPU and glibc
    transcendentals differ by ulps)."""
    N, T = 5000, 40
    obs = _obs(T)
    say("filter fp64", scheme)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.FP64_FUSED, scheme=scheme, seed=42,
                                resample_seed=4242))
    f.set_obs(obs)
    f.run(T)
    f.finish()
    h = f.history()
    st, lw = f.state()
    _, olw, rec = oracle.cv_run(N, obs, 42, 4242, scheme, 0.5)
    assert (h["resampled"] == rec["resampled"]).all()
    assert (h["draws"] == rec["draws"]).all()
    np.testing.assert_allclose(h["ess"], rec["ess"], rtol=1e-9)
    np.testing.assert_allclose(h["log_z"], rec["log_z"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(h["mean_x"], rec["mean_x"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(h["mean_y"], rec["mean_y"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)
    f.close()


def _tile_surplus(rep, tile=2048):
    """Surplus copies (rep - 1 over elements with rep >= 2) of every 2048-slot tile."""
    rep = np.asarray(rep, np.int64)
    s = np.maximum(rep - 1, 0)
    n = rep.size
    return np.add.reduceat(s, np.arange(0, n, tile))


@pytest.mark.parametrize("scheme", [V.SYSTEMATIC, V.STRATIFIED, V.MULTINOMIAL, V.RESIDUAL])
def test_filter_fp64_tiles_with_more_copies_than_slots(ctx, oracle, scheme):
    """A first observation at the edge of the prior makes a few particles
    carry almost all the weight (ESS about 1), so their tiles hold more
    surplus copies than the layout pass can stage in LDS: each copy then
    finds its element by binary search over the tile's running copy counts.
    The FP64 fused filter still equals the oracle's run of vSMC's algorithm
    (decisions and draws equal, values to 1e-9); the test first checks, with
    the oracle's own resampler on the oracle's first weights, that such tiles
    occur."""
    N, T = 1 << 16, 5
    obs = _obs(T).copy()
    obs[0] = (6.0, -6.0)  # three prior sds away in both coordinates
    _, lw0, _ = oracle.cv_run(N, obs[:1], 42, 4242, scheme, 0.0)
    w0 = np.exp(lw0 - lw0.max())
    w0 /= w0.sum()
    rep, _, _ = oracle.resample(V.SYSTEMATIC, w0, 4242)
    assert _tile_surplus(rep).max() > 2304  # the LDS staging's capacity
    say("filter fp64 overflow tiles", scheme)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.FP64_FUSED, scheme=scheme, seed=42,
                                resample_seed=4242))
    f.set_obs(obs)
    f.run(T)
    f.finish()
    h = f.history()
    st, lw = f.state()
    ost, olw, rec = oracle.cv_run(N, obs, 42, 4242, scheme, 0.5)
    assert h["resampled"][0] == 1
    assert (h["resampled"] == rec["resampled"]).all()
    assert (h["draws"] == rec["draws"]).all()
    np.testing.assert_allclose(h["ess"], rec["ess"], rtol=1e-9)
    np.testing.assert_allclose(h["log_z"], rec["log_z"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(h["mean_x"], rec["mean_x"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)
    f.close()


@pytest.mark.parametrize("scheme", range(6))
def test_filter_debug_step_ancestors(ctx, oracle, scheme):
    """Given the weights the device resampled from (and, for the sorted-uniform
    schemes, the uniforms it inverted), vSMC's sequential inversion and
    cfrp_trans give exactly the device's replication and ancestors."""
    _debug_step_ancestors(ctx, oracle, scheme, 20000, 6)


@pytest.mark.parametrize("scheme", range(6))
def test_filter_debug_step_ancestors_full_size(ctx, oracle, scheme):
    """The same at the bench's N = 2^24 (8192 fold tiles), two resampling
    steps: the second gathers through the first one's ancestors."""
    _debug_step_ancestors(ctx, oracle, scheme, 1 << 24, 2)


def _debug_step_ancestors(ctx, oracle, scheme, N, T):
    obs = _obs(T, 11)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.MIXED, scheme=scheme, threshold=1.01))
    f.set_obs(obs)
    for t in range(T):
        say("debug step", scheme, t)
        d = f.debug_step()
        assert d["resampled"]
        w = d["weight"]
        e0 = d["engine"]
        from oracle_ctypes import TF64Engine
        oeng = TF64Engine()
        for j in range(4):
            oeng.ctr[j] = e0.ctr[j]
            oeng.key[j] = e0.key[j]
            oeng.buf[j] = e0.buffer[j]
        oeng.index = e0.index
        if scheme in (V.MULTINOMIAL, V.RESIDUAL):
            if scheme == V.MULTINOMIAL:
                rep, _ = oracle.inversion_u(w, d["uniforms"])
            else:
                ip, rr, nn = oracle.residual_split(w)
                assert nn == d["nn"]
                rep, _ = oracle.inversion_u(rr, d["uniforms"])
                rep = rep + ip
        else:
            rep, _, _ = oracle.resample(scheme, w, engine=oeng)
        assert (d["rep"] == rep).all(), (scheme, t)
        assert (d["copy_from"] == oracle.cfrp_trans(rep)).all(), (scheme, t)
    f.close()


@pytest.mark.parametrize("scheme", [V.MULTINOMIAL, V.RESIDUAL])
def test_filter_ancestors_from_seed_full_size(ctx, oracle, scheme):
    """The sorted-uniform schemes at N = 2^24 from the seed: from the weights
    the device resampled and the engine state before the resample, the oracle
    runs vSMC's whole resampler (its own uniforms: the sequential FP64 fold
    lmax += log(1 - U) / (nn - n) of rng/u01.hpp:185-187 with glibc log and
    exp; then inversion and cfrp_trans).  The device evaluates the same fold
    exactly, so the uniforms can differ only where its log or exp rounds
    otherwise than glibc's: by at most one unit in the last place (about 4%
    of them).  The replication counts and ancestors are equal over three
    resampling steps: a one-ulp shift moves an ancestor only when a
    cumulative weight falls between the two neighbouring doubles."""
    N, T = 1 << 24, 3
    obs = _obs(T, 11)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.MIXED, scheme=scheme, threshold=1.01))
    f.set_obs(obs)
    from oracle_ctypes import TF64Engine
    for t in range(T):
        say("from seed", scheme, t)
        d = f.debug_step()
        assert d["resampled"]
        e0 = d["engine"]

        def engine():
            oe = TF64Engine()
            for j in range(4):
                oe.ctr[j] = e0.ctr[j]
                oe.key[j] = e0.key[j]
                oe.buf[j] = e0.buffer[j]
            oe.index = e0.index
            return oe

        w = d["weight"]
        nn = int(d["nn"])
        u_dev = d["uniforms"][:nn]
        u_ref, _ = oracle.u01seq(0, nn, engine=engine())
        du = np.flatnonzero(u_dev != u_ref)
        rep, _, _ = oracle.resample(scheme, w, engine=engine())
        bad = np.flatnonzero(d["rep"] != rep)
        say("from seed", scheme, t, "nn", nn, "uniforms differing", du.size,
            "max |du|", float(np.abs(u_dev - u_ref).max()), "replication differing", bad.size)
        assert bad.size == 0, (t, bad[:5])
        assert (d["copy_from"] == oracle.cfrp_trans(rep)).all(), t
        # one ulp of a value in [0.5, 1) at most
        assert float(np.abs(u_dev - u_ref).max()) <= 2.0 ** -53, t
        assert du.size <= nn // 10, du.size
    f.close()


def _rep_from_prefix(acc, v, nn):
    """vSMC's inversion (resample/common.hpp:74-83) from a given array of
    cumulative weights acc and the running max v of the uniforms: element i <
    M-1 takes the uniforms with acc_{i-1} < v_n <= acc_i, the last one the rest."""
    M = acc.size
    c = np.searchsorted(v, acc[:M - 1], side="right").astype(np.int64)
    c = np.maximum.accumulate(c)  # a scan that is not monotone takes nothing where it falls
    rep = np.empty(M, np.int64)
    rep[0] = c[0]
    rep[1:M - 1] = np.diff(c)
    rep[M - 1] = nn - c[-1]
    return rep.astype(np.uint64)


def _blocked_scan(w, dtype):
    """Inclusive prefix sums of w in dtype, rounded in the order of a one-pass
    blocked GPU scan: tiles of 2048 as 256 threads of 8 consecutive values; a
    sequential sum within each thread; a Hillis-Steele scan of the 256 thread
    sums; the tile totals chained in tile order (a decoupled look-back adds the
    previous tiles' prefix to each tile's total); then value i is (tile prefix
    + thread prefix) + the thread's own running sum."""
    n = w.size
    T = -(-n // 2048)
    x = np.zeros(T * 2048, dtype)
    x[:n] = w.astype(dtype)
    run = np.cumsum(x.reshape(T, 256, 8), axis=2, dtype=dtype)
    s = run[:, :, -1].copy()
    d = 1
    while d < 256:
        t = s.copy()
        t[:, d:] = s[:, d:] + s[:, :-d]
        s = t
        d *= 2
    thr = np.zeros_like(s)
    thr[:, 1:] = s[:, :-1]
    tile = np.zeros(T, dtype)
    tile[1:] = np.cumsum(s[:-1, -1], dtype=dtype)
    acc = (tile[:, None, None] + thr[:, :, None]) + run
    return acc.reshape(-1)[:n]


@pytest.mark.parametrize("scheme", [V.SYSTEMATIC, V.STRATIFIED, V.MULTINOMIAL])
def test_scan_precision_mismatch_full_size(ctx, oracle, scheme):
    """How many ancestors a cumulative-weight scan other than vSMC's sequential
    FP64 left fold misplaces at N = 2^24, on the filter's own weights and
    vSMC's uniforms from the engine state before the resample, over two
    resampling steps.  Scans: the rounding order of a blocked GPU scan
    (_blocked_scan) in FP64 and in FP32, and the sequential FP32 fold.  The
    exact fold (the device's own replication counts) misplaces none; an FP32
    scan misplaces a large share of them: at 2^24 one FP32 unit at the top of
    [0.5, 1) is the grid spacing 1/N itself."""
    N, T = 1 << 24, 2
    kind = {V.SYSTEMATIC: 2, V.STRATIFIED: 1, V.MULTINOMIAL: 0}[scheme]
    obs = _obs(T, 11)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.MIXED, scheme=scheme, threshold=1.01))
    f.set_obs(obs)
    from oracle_ctypes import TF64Engine
    for t in range(T):
        say("scan precision", scheme, t)
        d = f.debug_step()
        assert d["resampled"]
        e0 = d["engine"]
        oe = TF64Engine()
        for j in range(4):
            oe.ctr[j] = e0.ctr[j]
            oe.key[j] = e0.key[j]
            oe.buf[j] = e0.buffer[j]
        oe.index = e0.index
        u, _ = oracle.u01seq(kind, N, engine=oe)
        v = np.maximum.accumulate(u)
        w = d["weight"]
        acc = np.cumsum(w)  # numpy accumulates in order: vSMC's accw += w[i]
        rep = _rep_from_prefix(acc, v, N)
        assert (rep == d["rep"]).all(), (scheme, t)
        cf = oracle.cfrp_trans(rep)
        scans = {
            "fp64 blocked scan": _blocked_scan(w, np.float64),
            "fp32 blocked scan": _blocked_scan(w, np.float32).astype(np.float64),
            "fp32 sequential": np.cumsum(w.astype(np.float32), dtype=np.float32).astype(np.float64),
        }
        counts = {}
        for name, a in scans.items():
            r = _rep_from_prefix(a, v, N)
            counts[name] = (int((r != rep).sum()), int((oracle.cfrp_trans(r) != cf).sum()))
        say("scan precision", scheme, t, "N", N, "exact fold: 0 replication counts, 0 ancestors",
            "; ".join("%s: %d replication counts, %d ancestors" % (k, c[0], c[1])
                      for k, c in counts.items()))
        assert counts["fp32 blocked scan"][1] > N // 1000
        assert counts["fp32 sequential"][1] > N // 1000
    f.close()


def test_filter_mixed_first_step_coupled(ctx, oracle):
    """Before any resample has moved particles between slots the mixed pipeline
    sees the same particles as FP64: ESS within 1e-5 relative and log-evidence
    within 1e-5 absolute of the FP64 oracle (FP32 state and likelihood)."""
    N, T = 10000, 1
    obs = _obs(T, 5)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.MIXED, scheme=V.SYSTEMATIC))
    f.set_obs(obs)
    f.run(T)
    f.finish()
    h = f.history()
    _, _, rec = oracle.cv_run(N, obs, 42, 4242, V.SYSTEMATIC, 0.5)
    assert abs(h["ess"][0] - rec["ess"][0]) <= 1e-5 * rec["ess"][0]
    assert abs(h["log_z"][0] - rec["log_z"][0]) <= 1e-5
    f.close()


@pytest.mark.parametrize("N,T,scheme", [(1 << 16, 30, V.SYSTEMATIC), (1 << 16, 30, V.STRATIFIED),
                                         (1 << 24, 10, V.SYSTEMATIC), (1 << 24, 10, V.STRATIFIED),
                                         (1 << 24, 10, V.RESIDUAL), (1 << 24, 10, V.MULTINOMIAL),
                                         (1 << 24, 10, V.RESIDUAL_STRATIFIED),
                                         (1 << 24, 10, V.RESIDUAL_SYSTEMATIC)])
def test_filter_mixed_log_evidence_coupled(ctx, oracle, N, T, scheme):
    """The mixed pipeline's log-evidence against FP64 arithmetic on the mixed
    pipeline's own particles.  Step 0: the FP64 likelihood at the mixed prior
    draws (the same draws through vsmcx_state_cv_move, whose increments give
    the filter's first log-evidence exactly).  Step t > 0: the FP64 oracle
    re-runs the step from the mixed filter's population and weights (FP64
    normals, transition and likelihood).  Each step adds
    log sum_i W_i exp(inc_i).  Along the trajectory, resamples included, the
    mixed log-evidence stays within 1e-6 relative of that FP64 value, and
    within 1e-7 per step absolute on average: the error the FP32 state and
    likelihood add."""
    obs = _obs(T, 6)
    desc = V.cv_desc(N, T, precision=V.MIXED, scheme=scheme, seed=42, resample_seed=4242)
    s = V.State(ctx, N, 4, V.MIXED)
    inc32 = s.cv_move(desc, 0, obs[0])
    x0 = s.read()
    s.close()
    f = V.Filter(ctx, desc)
    f.set_obs(obs)
    f.run(1)

    def lme(a):
        m = a.max()
        return m + np.log(np.exp(a - m).sum())

    z0 = f.history(1)["log_z"][0]
    assert abs(z0 - (lme(inc32) - np.log(N))) <= 1e-12 * abs(z0)
    lz = lme(oracle.cv_llh_many(x0, obs[0, 0], obs[0, 1])) - np.log(N)
    worst = abs(z0 - lz) / abs(lz)
    for t in range(1, T):
        st_prev, lw = f.state()
        f.run(1)
        _, inc = oracle.cv_move_many(42, t, st_prev, obs[t, 0], obs[t, 1])
        lz += lme(lw + inc) - lme(lw)
        got = f.history(t + 1)["log_z"][t]
        worst = max(worst, abs(got - lz) / abs(lz))
    f.close()
    say("mixed coupled log-evidence N", N, "T", T, "final", got, "fp64", lz,
        "worst relative error", worst)
    assert worst <= 1e-6
    assert abs(got - lz) <= 1e-7 * T


def test_filter_mixed_vs_fp64_statistical(ctx, oracle):
    """After the first ancestor flip the mixed and FP64 runs are different Monte
    Carlo realisations (vSMC gives each slot its own stream, and one flip shifts
    cfrp_trans's zero-slot pairing).  Over 8 seeds the mean log-evidence and the
    mean final filter mean of the two pipelines agree within 4 standard errors."""
    N, T, R = 10000, 50, 8
    obs = _obs(T, 5)
    lz = {V.MIXED: [], V.FP64_FUSED: []}
    mx = {V.MIXED: [], V.FP64_FUSED: []}
    for r in range(R):
        for prec in lz:
            f = V.Filter(ctx, V.cv_desc(N, T, precision=prec, scheme=V.SYSTEMATIC, seed=1000 + r,
                                        resample_seed=77 + r))
            f.set_obs(obs)
            f.run(T)
            f.finish()
            h = f.history()
            lz[prec].append(h["log_z"][-1])
            mx[prec].append(h["mean_x"][-1])
            f.close()
    for d in (lz, mx):
        a, b = np.array(d[V.MIXED]), np.array(d[V.FP64_FUSED])
        se = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
        say("mixed vs fp64", a.mean(), b.mean(), se)
        assert abs(a.mean() - b.mean()) <= 4 * se + 1e-9


@pytest.mark.parametrize("scheme", [V.SYSTEMATIC, V.RESIDUAL])
def test_filter_plain_matches_oracle(ctx, oracle, scheme):
    """The un-fused FP64 pipeline (one kernel per vSMC operation) against the
    oracle: identical resampling decisions and draws, rtol 1e-9 elsewhere."""
    N, T = 5000, 40
    obs = _obs(T)
    f = V.Filter(ctx, V.cv_desc(N, T, precision=V.FP64_PLAIN, scheme=scheme))
    f.set_obs(obs)
    f.run(T)
    f.finish()
    h = f.history()
    st, lw = f.state()
    ost, olw, rec = oracle.cv_run(N, obs, 42, 4242, scheme, 0.5)
    assert (h["resampled"] == rec["resampled"]).all()
    assert (h["draws"] == rec["draws"]).all()
    np.testing.assert_allclose(h["ess"], rec["ess"], rtol=1e-9)
    np.testing.assert_allclose(h["log_z"], rec["log_z"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(h["mean_x"], rec["mean_x"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st, ost, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lw, olw, rtol=1e-9, atol=1e-9)
    f.close()


def test_degenerate_weights_error(ctx):
    """All-zero weights are rejected (vSMC would produce NaN silently)."""
    eng = V.ThreefryState.seeded(1)
    with pytest.raises(V.VsmcxError):
        ctx.resample(V.SYSTEMATIC, np.array([1.0, -1.0]), eng)
