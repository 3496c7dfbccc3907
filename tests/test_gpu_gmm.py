"""Gaussian-mixture SMC with adaptive tempering (vsmcx_gmm_*) on the GPU,
through the C ABI.

  * FP64: every iteration's record (alpha, alpha_inc, ESS, resampling
    decision, bisection steps, acceptance counts, SMC and path-sampling log
    evidence) and the final particles against the CPU oracle
    (oracle/gmm_oracle.c, itself pinned by quadrature and prior sampling in
    tests/test_gmm_oracle.py) on the same seeds: decisions and counts equal,
    values within 1e-9 relative;
  * mixed precision (FP32 parameters and likelihood terms, FP64 weights,
    search and evidence): the initial log likelihoods within 1e-5 relative of
    FP64, and the final log evidence statistically equal to FP64 over seeds;
  * argument checks.
"""
import numpy as np
import pytest

import vsmc_amd as V

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = V.Context(0)
    yield c
    c.close()


def _run(ctx, y, n, K, sweeps, precision, seed, rseed, scheme=V.STRATIFIED):
    g = V.Gmm(ctx, y, n, components=K, sweeps=sweeps, precision=precision, seed=seed,
              resample_seed=rseed, scheme=scheme)
    g.initialize()
    g.iterate()
    h = g.history()
    st = g.state()
    g.close()
    return h, st


@pytest.mark.parametrize("K,ndata,n,sweeps,scheme", [(4, 200, 2048, 2, V.STRATIFIED),
                                                     (2, 60, 3000, 1, V.SYSTEMATIC),
                                                     (6, 80, 1500, 1, V.STRATIFIED)])
def test_gmm_fp64_vs_oracle(ctx, oracle, K, ndata, n, sweeps, scheme):
    y = V.gmm_simulate(ndata, 5)
    say("gmm fp64", K, ndata, n, sweeps, scheme)
    h, (p, llh, lw) = _run(ctx, y, n, K, sweeps, V.FP64_FUSED, 21, 22, scheme)
    op, ollh, olw, rec = oracle.gmm_run(n, y, K=K, sweeps=sweeps, seed=21, resample_seed=22,
                                        scheme=scheme, threads=16)
    assert len(h) == len(rec)
    assert h["alpha"][-1] == 1.0
    for f in ("resampled", "bisections"):
        assert (h[f] == rec[f]).all(), f
    assert (h["accept"] == rec["accept"]).all()
    for f in ("alpha", "alpha_inc", "ess", "log_z", "path_integrand", "path_log_z"):
        np.testing.assert_allclose(h[f], rec[f], rtol=1e-9, atol=1e-12, err_msg=f)
    np.testing.assert_allclose(p, op, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(llh, ollh, rtol=1e-9)
    w, ow = np.exp(lw - lw.max()), np.exp(olw - olw.max())
    np.testing.assert_allclose(w / w.sum(), ow / ow.sum(), rtol=1e-9, atol=1e-15)


def test_gmm_mixed_initial_likelihoods(ctx, oracle):
    y = V.gmm_simulate(1000, 7)
    n = 4096
    g64 = V.Gmm(ctx, y, n, precision=V.FP64_FUSED, seed=3)
    g32 = V.Gmm(ctx, y, n, precision=V.MIXED, seed=3)
    g64.initialize()
    g32.initialize()
    p64, l64, _ = g64.state()
    p32, l32, _ = g32.state()
    g64.close()
    g32.close()
    # The same Philox words; the FP32 path converts them with 23-bit midpoint
    # uniforms (cells of 2^-23), so an exponential -log(u) moves by more than
    # 2e-5 relative wherever 1 - u < ~3e-3: about 0.15% of the particles.
    # Everywhere else the draws agree to FP32 rounding.
    same = np.isclose(p32, p64, rtol=2e-5, atol=2e-5).all(1)
    assert same.mean() > 0.995
    # The FP32 likelihood of the FP32 parameters against the FP64 likelihood
    # of the same parameters (the likelihood moves by ~1e4 per unit lambda
    # for prior draws, so the rounding of the parameters themselves is not
    # compared here).
    y32 = y.astype(np.float32).astype(np.float64)
    for i in range(0, n, 16):
        want = oracle.gmm_log_likelihood(4, p32[i], y32)
        assert abs(l32[i] - want) <= 1e-5 * abs(want), (i, l32[i], want)
    np.testing.assert_allclose(l64, [oracle.gmm_log_likelihood(4, p64[i], y) for i in range(n)],
                               rtol=1e-12)


def test_gmm_mixed_vs_fp64_log_evidence(ctx):
    y = V.gmm_simulate(300, 9)
    n, R = 1 << 14, 4
    z = {}
    steps = {}
    for prec in (V.FP64_FUSED, V.MIXED):
        zs, ns = [], []
        for r in range(R):
            h, _ = _run(ctx, y, n, 4, 2, prec, 100 + r, 200 + r)
            zs.append(h["log_z"][-1])
            ns.append(len(h))
        z[prec] = np.array(zs)
        steps[prec] = np.mean(ns)
        say("gmm", prec, z[prec], ns)
    se = np.hypot(z[V.MIXED].std(ddof=1), z[V.FP64_FUSED].std(ddof=1)) / np.sqrt(R)
    assert abs(z[V.MIXED].mean() - z[V.FP64_FUSED].mean()) < 4 * se + 0.05
    assert abs(steps[V.MIXED] - steps[V.FP64_FUSED]) <= 0.2 * steps[V.FP64_FUSED]


def test_gmm_rejects_bad_arguments(ctx):
    y = V.gmm_simulate(50, 1)
    with pytest.raises(V.VsmcxError):
        V.Gmm(ctx, y, 100, components=7)
    with pytest.raises(V.VsmcxError):
        V.Gmm(ctx, np.ones(10), 100)
    with pytest.raises(V.VsmcxError):
        V.Gmm(ctx, y, 100, cess_drop=1.5)
    g = V.Gmm(ctx, y, 100)
    with pytest.raises(V.VsmcxError):
        g.iterate()  # not initialised
    g.close()
