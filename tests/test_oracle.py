"""Pins the CPU oracle (oracle/vsmc_oracle.c) before anything is checked against it.

Sources of truth, strongest first:
  * golden vectors written by the compiled vSMC headers (tests/golden/),
  * Random123 published known-answer vectors for the engine block functions,
  * the live compiled headers (oracle/_ref) when present in this container,
  * closed forms, invariants and brute force on tiny inputs.
"""
import ctypes as C
import itertools
from fractions import Fraction

import numpy as np
import pytest

from oracle_ctypes import (RESIDUAL, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC, STRATIFIED,
                           SYSTEMATIC, MULTINOMIAL, RefHarness, ref_available)

# Random123 kat_vectors (SC11 distribution): philox4x32_10, threefry4x64_20,
# threefry4x32_20 on the all-zero, all-ones and pi-digit inputs.
KAT_PHILOX = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]
KAT_TF64 = [
    ([0] * 4, [0] * 4,
     [0x09218ebde6c85537, 0x55941f5266d86105, 0x4bd25e16282434dc, 0xee29ec846bd2e40b]),
    ([2**64 - 1] * 4, [2**64 - 1] * 4,
     [0x29c24097942bba1b, 0x0371bbfb0f6f4e11, 0x3c231ffa33f83a1c, 0xcd29113fde32d168]),
]
KAT_TF32 = [
    ([0] * 4, [0] * 4, [0x9c6ca96a, 0xe17eae66, 0xfc10ecd4, 0x5256a7d8]),
    ([0xffffffff] * 4, [0xffffffff] * 4, [0x2a881696, 0x57012287, 0xf6c7446e, 0xa16a6732]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
     [0xa4093822, 0x299f31d0, 0x082efa98, 0xec4e6c89],
     [0x59cd1dbb, 0xb8879579, 0x86b5d00c, 0xac8b6d84]),
]


@pytest.mark.parametrize("ctr,key,want", KAT_PHILOX)
def test_philox_kat(oracle, ctr, key, want):
    assert list(oracle.philox4x32_block(ctr, key)) == want


@pytest.mark.parametrize("ctr,key,want", KAT_TF64)
def test_threefry4x64_kat(oracle, ctr, key, want):
    assert [int(x) for x in oracle.threefry4x64_block(ctr, key)] == want


@pytest.mark.parametrize("ctr,key,want", KAT_TF32)
def test_threefry4x32_kat(oracle, ctr, key, want):
    assert list(oracle.threefry4x32_block(ctr, key)) == want


def test_blocks_golden(oracle, golden):
    for c, k, w in zip(golden["philox_ctr"], golden["philox_key"], golden["philox_out"]):
        assert (oracle.philox4x32_block(c, k) == w).all()
    for c, k, w in zip(golden["tf64_ctr"], golden["tf64_key"], golden["tf64_out"]):
        assert (oracle.threefry4x64_block(c, k) == w).all()
    for c, k, w in zip(golden["tf32_ctr"], golden["tf32_key"], golden["tf32_out"]):
        assert (oracle.threefry4x32_block(c, k) == w).all()


def test_engine_streams_golden(oracle, golden):
    """Increment-before-generate, word order 0..3, seed -> key[0], discard, carries."""
    for s, ph, tf in zip(golden["stream_skips"], golden["ph32_streams"], golden["tf64_streams"]):
        assert (oracle.ph32_stream(2024, 23, int(s)) == ph).all()
        assert (oracle.tf64_stream(2024, 23, int(s)) == tf).all()
    assert (oracle.tf32_stream(2024, 23) == golden["tf32_stream"]).all()
    assert (oracle.ph32_stream(5, 23, 0, [2**32 - 1, 2**32 - 1, 7, 0]) == golden["ph32_carry"]).all()
    assert (oracle.tf64_stream(5, 23, 0, [2**64 - 2, 2**64 - 1, 5, 0]) == golden["tf64_carry"]).all()


def test_first_block_uses_ctr_one(oracle):
    """Engine seeded with s draws block ctr={1,0,0,0} first (philox.hpp:386-395)."""
    assert (oracle.ph32_stream(9, 4) == oracle.philox4x32_block([1, 0, 0, 0], [9, 0])).all()
    assert (oracle.tf64_stream(9, 4) == oracle.threefry4x64_block([1, 0, 0, 0], [9, 0, 0, 0])).all()


def test_canonical_golden(oracle, golden):
    assert (oracle.canonical_tf64(101, 256) == golden["canonical_tf64"]).all()
    assert (oracle.canonical_ph32(101, 256) == golden["canonical_ph32"]).all()


def test_canonical_closed_form(oracle):
    """64-bit engine: U = double(x) / 2^64 with the x -> 2^64 rounding clamp."""
    xs = oracle.tf64_stream(101, 64)
    us = oracle.canonical_tf64(101, 64)
    for x, u in zip(xs, us):
        want = float(int(x)) / 2.0**64
        if want >= 1.0:
            want = np.nextafter(1.0, 0.0)
        assert u == want


@pytest.mark.parametrize("kind,name", [(0, "sorted"), (1, "stratified"), (2, "systematic")])
def test_u01seq_golden(oracle, golden, kind, name):
    u, e = oracle.u01seq(kind, 1000, 77)
    assert (u == golden[f"u01seq_{name}"]).all()
    nxt = np.zeros(4, np.uint64)
    oracle.lib.orc_tf64_stream(C.byref(e), 4, nxt)
    assert (nxt == golden[f"u01seq_{name}_next"]).all()


def test_resample_golden(oracle, golden):
    n = int(golden["resample_ncases"])
    for ci in range(n):
        w = golden[f"resample_w_{ci}"]
        for scheme in range(6):
            rep, draws, e = oracle.resample(scheme, w, 1000 + ci)
            assert (rep == golden[f"resample_rep_{ci}_{scheme}"]).all(), (ci, scheme)
            assert (oracle.cfrp_trans(rep) == golden[f"resample_cf_{ci}_{scheme}"]).all()
            nxt = np.zeros(4, np.uint64)
            oracle.lib.orc_tf64_stream(C.byref(e), 4, nxt)
            assert (nxt == golden[f"resample_next_{ci}_{scheme}"]).all(), (ci, scheme, draws)


def test_weightset_golden(oracle, golden):
    lw, w, ess = oracle.weightset(golden["ws_lw0"])
    assert (lw == golden["ws_set_lw"]).all() and (w == golden["ws_set_w"]).all()
    assert ess == golden["ws_set_ess"]
    lw, w, ess = oracle.weightset(golden["ws_lw0"], golden["ws_inc"])
    assert (lw == golden["ws_add_lw"]).all() and (w == golden["ws_add_w"]).all()
    assert ess == golden["ws_add_ess"]


@pytest.mark.parametrize("scheme", range(6))
def test_cv_run_golden(oracle, golden, scheme):
    """Whole filter (do_init/do_iter replay) equals the Particle<T>-driven run bit for bit."""
    st, lw, rec = oracle.cv_run(1000, golden["cv_obs"], 42, 4242, scheme, 0.5)
    for f in rec.dtype.names:
        assert (rec[f] == golden[f"cv_{scheme}_{f}"]).all(), f
    assert (st == golden[f"cv_{scheme}_state"]).all()
    assert (lw == golden[f"cv_{scheme}_lw"]).all()


# ---------------- closed forms / invariants / brute force ----------------

def test_fold_is_left_fold(oracle):
    """1 + 2^-53 + 2^-53 folds to 1 (ties-to-even twice); a pairwise sum gives 1+2^-52."""
    e = 2.0**-53
    assert oracle.fold_prefix([1.0, e, e])[-1] == 1.0
    assert oracle.fold_prefix([e, e, 1.0])[-1] == 1.0 + 2 * e
    x = np.array([0.1] * 10)
    assert oracle.fold_prefix(x)[-1] == 0.9999999999999999


def _brute_inversion(w, u):
    """rep_i = #{n : bin(u_n) = i}, bin = running max of first i with u <= cw_i (exact)."""
    M = len(w)
    cw, a = [], Fraction(0)
    for x in w[:-1]:
        a += x
        cw.append(a)
    rep = [0] * M
    b = 0
    for un in u:
        while b < M - 1 and not (un <= cw[b]):
            b += 1
        rep[b] += 1
    return rep


def test_inversion_bruteforce_dyadic(oracle):
    """Dyadic weights/uniforms make every FP64 sum exact, so the exact-rational
    brute force must agree with the oracle's FP64 inversion everywhere."""
    for M, N in itertools.product(range(1, 6), range(0, 6)):
        for trial in range(20):
            rng = np.random.default_rng(M * 100 + N * 10 + trial)
            k = rng.integers(0, 9, M).astype(float)
            if k.sum() == 0:
                k[0] = 1
            w = k / 8.0 / (k.sum() / 8.0)
            w = np.round(w * 1024) / 1024
            w[-1] = 1.0 - w[:-1].sum()
            if (w < 0).any():
                continue
            u = np.sort(np.round(rng.random(N) * 256) / 256)
            rep, _ = oracle.inversion_u(w, u)
            want = _brute_inversion([Fraction(x) for x in w], [Fraction(x) for x in u])
            assert list(rep) == want, (M, N, w, u)


@pytest.mark.parametrize("scheme", range(6))
def test_resample_invariants(oracle, scheme):
    rng = np.random.default_rng(scheme)
    for N in (1, 2, 3, 10, 257, 4096):
        w = rng.random(N) ** 4
        w[rng.integers(0, N, N // 3)] = 0
        if w.sum() == 0:
            w[0] = 1
        w /= w.sum()
        rep, draws, _ = oracle.resample(scheme, w, 5)
        rep = rep.astype(np.int64)
        assert rep.sum() == N
        assert (rep[:-1][w[:-1] == 0] == 0).all()
        if scheme == SYSTEMATIC:
            assert (np.abs(rep - N * w) < 1 + 1e-9).all() and draws == 1
        if scheme == STRATIFIED:
            assert (np.abs(rep - N * w) < 2 + 1e-9).all()
        if scheme in (RESIDUAL, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC):
            assert (rep >= np.floor(N * w) - 1e-9).all()
        cf = oracle.cfrp_trans(rep.astype(np.uint64)).astype(np.int64)
        keep = rep > 0
        assert (cf[keep] == np.arange(N)[keep]).all()
        assert (rep[cf[~keep]] >= 2).all()
        # gather is a permutation-with-replacement matching the replication counts
        assert (np.bincount(cf, minlength=N) == rep).all()


def test_multinomial_draw_count(oracle):
    """Sorted/stratified sequences draw lazily: min(N, n_final + 1) uniforms."""
    w = np.zeros(8)
    w[-1] = 1.0  # every uniform lands in the last bin: inversion reads u_0 only
    rep, draws, _ = oracle.resample(MULTINOMIAL, w, 3)
    assert list(rep) == [0] * 7 + [8] and draws == 1
    w = np.full(8, 1 / 8)
    rep, draws, _ = oracle.resample(STRATIFIED, w, 3)
    assert rep.sum() == 8 and draws <= 8


def test_weightset_closed_form(oracle):
    lw, w, ess = oracle.weightset(np.zeros(64))
    assert ess == 64.0 and (w == 1 / 64).all() and (lw == 0).all()
    rng = np.random.default_rng(0)
    lw, w, ess = oracle.weightset(rng.normal(size=1000) * 3)
    assert abs(w.sum() - 1) < 1e-13 and lw.max() == 0.0
    assert abs(ess - 1 / np.sum(w * w)) < 1e-9 * ess


def test_cv_normals_are_box_muller(oracle):
    z = oracle.cv_normals(42, 7, 3)
    blk = oracle.philox4x32_block([4, 7, 0, 0], [42, 0])
    u1, u2 = (1.0 + float(blk[0])) / 2**32, (1.0 + float(blk[1])) / 2**32
    r = np.sqrt(-2 * np.log(u1))
    assert abs(z[0] - r * np.cos(2 * np.pi * u2)) < 1e-12
    assert abs(z[1] - r * np.sin(2 * np.pi * u2)) < 1e-12


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (needs the vSMC tree)")
def test_oracle_vs_live_reference(oracle):
    r = RefHarness()
    rng = np.random.default_rng(99)
    for scheme in range(6):
        for N in (3, 64, 999, 5000):
            w = rng.random(N) ** 6
            w /= w.sum()
            rep, _, e = oracle.resample(scheme, w, 11)
            rrep, rcf, nxt = r.resample(scheme, w, 11)
            assert (rep == rrep).all() and (oracle.cfrp_trans(rep) == rcf).all()
            onxt = np.zeros(4, np.uint64)
            oracle.lib.orc_tf64_stream(C.byref(e), 4, onxt)
            assert (onxt == nxt).all()


def test_cv_move_many_is_cv_run_step(oracle):
    """The vectorised step used by the coupled mixed-precision test is the
    oracle filter's own step: the same states and increments after step 0, and
    after a second step with no resample in between."""
    import vsmc_amd as V
    N = 300
    obs = V.cv_simulate(2, 4)
    st1, lw1, rec = oracle.cv_run(N, obs[:1], 42, 4242, 3, 0.0)
    s0, inc0 = oracle.cv_move_many(42, 0, np.zeros((N, 4)), obs[0, 0], obs[0, 1])
    assert (s0 == st1).all()
    assert (oracle.cv_llh_many(s0, obs[0, 0], obs[0, 1]) == inc0).all()
    st2, lw2, _ = oracle.cv_run(N, obs, 42, 4242, 3, 0.0)  # threshold 0: never resamples
    s1, inc1 = oracle.cv_move_many(42, 1, s0, obs[1, 0], obs[1, 1])
    assert (s1 == st2).all()
    lw = inc0 + inc1
    np.testing.assert_allclose(lw - lw.max(), lw2, rtol=0, atol=1e-12)


def test_cv_llh_is_the_student_t_density(oracle):
    """The CV model's log-likelihood increment is the log density of two
    independent Student-t(nu = 10, scale 0.1) coordinates, as scipy evaluates
    it.  The reference's pf likelihood (example/pf/include/pf_smp.hpp:96-111)
    is the same up to the per-coordinate constant lgamma((nu+1)/2) -
    lgamma(nu/2) - log(nu pi)/2 - log(scale) (llh_const), which shifts log Z by
    2 llh_const per step and no weight; this pins that constant and the
    oracle's whole expression independently of both."""
    from scipy import stats

    import vsmc_amd as V
    rng = np.random.default_rng(31)
    st = rng.normal(size=(2000, 4)) * np.array([0.3, 0.3, 1.0, 1.0])
    st[:10, :2] = 0.25  # at the observation: the density's peak
    for ox, oy in ((0.25, 0.25), (-1.3, 2.0)):
        got = oracle.cv_llh_many(st, ox, oy)
        want = (stats.t.logpdf(st[:, 0] - ox, df=10, scale=0.1) +
                stats.t.logpdf(st[:, 1] - oy, df=10, scale=0.1))
        assert np.allclose(got, want, rtol=1e-13, atol=1e-12)
    for nu, s in ((10.0, 0.1), (3.0, 2.5)):
        want = float(stats.t.logpdf(0.0, df=nu, scale=s))
        assert abs(V.cv_llh_const(nu, s) - want) < 1e-13
        from oracle_ctypes import cv_param
        assert abs(cv_param(nu=nu, obs_scale=s).llh_const - want) < 1e-13
