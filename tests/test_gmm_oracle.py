"""The GMM tempered-SMC oracle (oracle/gmm_oracle.c) pinned by what the
mathematics fixes, independently of its own formulas:

  * K = 1: the log evidence is a 2-D integral over (mu, log lambda), computed
    here by quadrature on a fine grid; the SMC estimate (mean of 8 runs) must
    sit within 4 standard errors.  A dropped lambda Jacobian, a reweight by
    alpha instead of alpha_inc, a flipped MH test or a wrong prior variance
    each move the estimate by 15 to 150 standard errors.
  * K = 2 (exercises the logit-weight move): the log evidence by importance
    sampling from the prior (4e6 draws with numpy's normal, gamma and
    Dirichlet generators); a dropped logit Jacobian moves the SMC estimate by
    about 80 standard errors.
  * the log likelihood against scipy's normal density; the initial draws
    against the moments of N(mu0, sd0^2), Gamma(2, scale0) and Dirichlet(1);
    the prior constants against their definition from the data range
    (example/gmm/include/gmm_state.hpp:145-160);
  * the tempering schedule: alpha increases to exactly 1, alpha_inc is the
    step in alpha, the resampling decision follows ESS < threshold * N.
"""
import os

import numpy as np
import pytest
from scipy import stats

LN2PI = np.log(2 * np.pi)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _k1_data():
    rng = np.random.default_rng(11)
    return 1.0 + 0.7 * rng.standard_normal(20)


def _k2_data():
    rng = np.random.default_rng(12)
    return np.concatenate([-1.5 + 0.6 * rng.standard_normal(5), 1.5 + 0.6 * rng.standard_normal(5)])


def _smc_logz(oracle, y, K, N, R):
    z = np.array([oracle.gmm_run(N, y, K=K, sweeps=3, seed=100 + s, resample_seed=200 + s,
                                 threads=8, want_state=False)[3][-1]["log_z"] for s in range(R)])
    return z.mean(), z.std(ddof=1) / np.sqrt(R)


def test_gmm_prior_constants(oracle):
    y = np.array([3.0, -2.0, 0.5, 7.0])
    mu0, sd0, shape0, scale0 = oracle.gmm_prior(y)
    kappa = 1 / 81.0
    assert mu0 == 2.5 and shape0 == 2
    assert np.isclose(sd0, 9.0) and np.isclose(scale0, 50 * kappa)


def test_gmm_log_likelihood_vs_scipy(oracle):
    rng = np.random.default_rng(3)
    y = rng.normal(size=50) * 3
    for K in (1, 3, 6):
        mu = rng.normal(size=K)
        lam = rng.gamma(2.0, 1.0, K)
        w = rng.dirichlet(np.ones(K))
        p = np.concatenate([mu, lam, w])
        dens = (w[None, :] * stats.norm.pdf(y[:, None], mu[None, :], 1 / np.sqrt(lam[None, :]))).sum(1)
        # the reference's likelihood omits 1/sqrt(2 pi) per point and adds 1e-13
        want = np.log(dens * np.sqrt(2 * np.pi) + 1e-13).sum()
        assert np.isclose(oracle.gmm_log_likelihood(K, p, y), want, rtol=1e-12)


def test_gmm_init_moments(oracle):
    prior = np.array([0.5, 2.0, 2.0, 0.3])
    K, n = 3, 60000
    ps = np.array([oracle.gmm_init_particle(K, 9, i, prior) for i in range(n)])
    mu, lam, w = ps[:, :K], ps[:, K:2 * K], ps[:, 2 * K:]
    se = 1 / np.sqrt(n)
    assert abs(mu.mean() - 0.5) < 5 * 2.0 * se and abs(mu.std() - 2.0) < 5 * 2.0 * se
    assert abs(lam.mean() - 0.6) < 5 * np.sqrt(2) * 0.3 * se      # Gamma(2, 0.3): mean 0.6
    assert abs(lam.var() - 2 * 0.09) < 0.01                        # variance 0.18
    assert np.allclose(w.sum(1), 1)
    assert abs(w.mean() - 1 / K) < 0.005 and abs(w[:, 0].var() - (K - 1) / (K * K * (K + 1))) < 0.003
    # independent streams per particle and per component
    assert abs(np.corrcoef(mu[:, 0], mu[:, 1])[0, 1]) < 0.02


def _logz_k1_quadrature(y, prior):
    mu0, sd0, sh, sc = prior
    mu = np.linspace(y.mean() - 3, y.mean() + 3, 2001)
    u = np.linspace(-6, 4, 2001)
    M, U = np.meshgrid(mu, u, indexing="ij")
    lam = np.exp(U)
    lpm = -0.5 * ((M - mu0) / sd0) ** 2 - np.log(sd0) - 0.5 * LN2PI
    lpl = (sh - 1) * U - lam / sc - sh * np.log(sc) + U        # Gamma(2, sc) in log lambda
    ll = np.zeros_like(M)
    for x in y:
        ll += np.log(np.sqrt(lam) * np.exp(-0.5 * lam * (x - M) ** 2) + 1e-13)
    f = lpm + lpl + ll
    m = f.max()
    inner = np.trapezoid(np.exp(f - m), u, axis=1)
    return np.log(np.trapezoid(inner, mu)) + m - 0.5 * len(y) * LN2PI


def test_gmm_oracle_k1_log_evidence_vs_quadrature(oracle):
    y = _k1_data()
    exact = _logz_k1_quadrature(y, oracle.gmm_prior(y))
    m, se = _smc_logz(oracle, y, 1, 16384, 8)
    assert abs(m - exact) < 4 * np.hypot(se, 1e-3), (m, se, exact)


def _logz_prior_is(oracle, y, K, draws, seed):
    rng = np.random.default_rng(seed)
    mu0, sd0, _, sc = oracle.gmm_prior(y)
    out = []
    for _ in range(draws // 500000):
        m = 500000
        mu = mu0 + sd0 * rng.standard_normal((m, K))
        lam = rng.gamma(2.0, sc, (m, K))
        w = rng.dirichlet(np.ones(K), m)
        ll = np.zeros(m)
        for x in y:
            ll += np.log((w * np.sqrt(lam) * np.exp(-0.5 * lam * (x - mu) ** 2)).sum(1) + 1e-13)
        out.append(ll)
    ll = np.concatenate(out)
    mx = ll.max()
    e = np.exp(ll - mx)
    return np.log(e.mean()) + mx - 0.5 * len(y) * LN2PI, e.std() / e.mean() / np.sqrt(e.size)


def test_gmm_oracle_k2_log_evidence_vs_prior_sampling(oracle):
    y = _k2_data()
    exact, se_is = _logz_prior_is(oracle, y, 2, 4000000, 1)
    m, se = _smc_logz(oracle, y, 2, 8192, 16)
    assert abs(m - exact) < 4 * np.hypot(se, se_is), (m, se, exact, se_is)


def test_gmm_oracle_schedule_invariants(oracle):
    rng = np.random.default_rng(5)
    comp = rng.integers(0, 4, 100)
    y = np.array([-3, 0, 2, 5.0])[comp] + rng.uniform(0.5, 1.5, 4)[comp] * rng.standard_normal(100)
    N = 2000
    p, llh, lw, rec = oracle.gmm_run(N, y, K=4, sweeps=2, seed=1, resample_seed=2, threads=8)
    a = rec["alpha"]
    assert a[0] == 0 and a[-1] == 1 and (np.diff(a) > 0).all()
    assert np.allclose(rec["alpha_inc"][1:], np.diff(a), rtol=0, atol=1e-15)
    assert (rec["bisections"][1:] >= 1).all()
    assert ((rec["ess"][1:] < 0.5 * N) == (rec["resampled"][1:] == 1)).all()
    assert (rec["ess"] <= N * (1 + 1e-12)).all() and (rec["accept"][1:] > 0).all()
    # the stored log likelihood is the likelihood of the stored parameters
    for i in (0, 17, N - 1):
        assert np.isclose(llh[i], oracle.gmm_log_likelihood(4, p[i], y), rtol=1e-13)
    # SMC and path-sampling estimates of the same quantity agree roughly
    assert abs(rec["log_z"][-1] - rec["path_log_z"][-1]) < 1.0


# ---------------------------------------------------------------- tempering search
# Golden vectors written by tests/golden/make_golden.py from vSMC's own
# WeightSet::cess (include/vsmc/core/weight_set.hpp:161-186, 411-431) and the
# gmm example's alpha_ess<T, cess01<T>>::alpha_iter
# (example/include/move_alpha.hpp:234-279) on a Particle<T> with the
# gmm_state::alpha clamp (example/gmm/include/gmm_state.hpp:163-172).
# 38 cases: N = 1 .. 20000, equal / spread / heavy log-weights, log-likelihood
# spreads from 0.5 to 3000 and two where every multiplier underflows (CESS
# 0/0 = NaN, which moves the search up), start temperatures 0 .. 1 - 1e-7 and
# drops 0.1, 0.5 and 1e-9 (the tolerance floor 1e-10).

def _temper_cases(golden):
    for ci in range(int(golden["temper_ncases"])):
        alpha, drop, a_diff = golden[f"temper_in_{ci}"]
        a, ai, calls, cess = golden[f"temper_out_{ci}"]
        yield (ci, golden[f"temper_lw_{ci}"], golden[f"temper_llh_{ci}"], golden[f"temper_w_{ci}"],
               alpha, drop, a_diff, a, ai, int(calls), cess)


def _temper_mismatches(o, golden):
    bad = []
    for ci, lw, llh, w, alpha, drop, a_diff, a, ai, calls, cess in _temper_cases(golden):
        got = o.gmm_alpha_search(w, llh, alpha, drop)
        c = o.gmm_cess(w, llh, a_diff)
        same_cess = (np.isnan(c) and np.isnan(cess)) or c == cess
        if got != (a, ai, calls) or not same_cess:
            bad.append((ci, got, (a, ai, calls), c, cess))
    return bad


def test_gmm_weights_equal_reference_weight_set(oracle, golden):
    """The normalised weights the search reads are the oracle's WeightSet's."""
    for ci, lw, _, w, *_ in _temper_cases(golden):
        _, ow, _ = oracle.weightset(lw)
        assert (ow == w).all(), ci


def test_gmm_cess_and_alpha_search_equal_reference(oracle, golden):
    """CESS values and the search's alpha, alpha_inc and number of CESS
    evaluations (the bisection count the records carry) equal vSMC's, bit for
    bit, in every case."""
    assert _temper_mismatches(oracle, golden) == []


def test_gmm_alpha_search_equals_live_reference(oracle):
    from oracle_ctypes import RefHarness, ref_available
    if not ref_available():
        pytest.skip("oracle/_ref not built (needs the vSMC tree)")
    r = RefHarness()
    rng = np.random.default_rng(77)
    for _ in range(40):
        N = int(rng.integers(1, 5000))
        lw = rng.normal(size=N) * rng.choice([0.0, 1.0, 10.0])
        llh = -rng.random() * 1e4 + rng.normal(size=N) * rng.choice([0.1, 10.0, 1000.0])
        alpha, drop = float(rng.random()), float(rng.choice([0.1, 0.05, 0.3]))
        a, ai, calls, w = r.alpha_ess(lw, llh, alpha, drop)
        assert oracle.gmm_alpha_search(w, llh, alpha, drop) == (a, ai, calls)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_tempering_mistakes_fail_the_golden_vectors(golden, tmp_path, plant):
    """A copy of the oracle built with a plausible mistake planted
    (ORC_PLANT_CESS: 1 an ESS target instead of the CESS, 2 use_log semantics,
    3 a wrong start of the search) disagrees with vSMC's vectors."""
    import subprocess

    from oracle_ctypes import Oracle
    src = os.path.join(ROOT, "oracle")
    so = str(tmp_path / "liboracle_planted.so")
    subprocess.run(["gcc", "-std=c99", "-O2", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared",
                    f"-DORC_PLANT_CESS={plant}", "-o", so, os.path.join(src, "vsmc_oracle.c"),
                    os.path.join(src, "gmm_oracle.c"), "-lm"], check=True)
    bad = _temper_mismatches(Oracle(so), golden)
    assert len(bad) >= 5, bad


# ---------------------------------------------------------------- path sampling
def _k1_tempered_grid(y, prior):
    """log prior and log likelihood (without the 2 pi constant) on a K = 1 grid
    that covers the whole prior (mu0 +- 9 sd0; log lambda from log(scale0) - 14
    to log(scale0) + 4.5), as E_alpha[log L] at small alpha needs."""
    mu0, sd0, sh, sc = prior
    mu = np.linspace(mu0 - 9 * sd0, mu0 + 9 * sd0, 3001)
    u = np.linspace(np.log(sc) - 14, np.log(sc) + 4.5, 1501)
    M, U = np.meshgrid(mu, u, indexing="ij")
    lam = np.exp(U)
    lpm = -0.5 * ((M - mu0) / sd0) ** 2 - np.log(sd0) - 0.5 * LN2PI
    lpl = (sh - 1) * U - lam / sc - sh * np.log(sc) + U
    ll = np.zeros_like(M)
    for x in y:
        ll += np.log(np.sqrt(lam) * np.exp(-0.5 * lam * (x - M) ** 2) + 1e-13)
    return lpm + lpl, ll, mu, u


def test_gmm_path_log_z_vs_quadrature(oracle):
    """The path-sampling estimate (trapezoid over the adaptive alpha grid of
    E_alpha[log L], include/vsmc/core/path.hpp:229-239) at K = 1.  For each
    run, E_alpha[log L] is computed by quadrature at that run's own alphas;
    its trapezoid is what the estimator targets on that grid, and the mean
    over 8 runs of (path_log_z - that trapezoid) must sit within 4 standard
    errors.  The alpha grid's own discretisation error is bounded, not
    assumed: d/dalpha E_alpha[log L] = Var_alpha[log L] >= 0, so the integrand
    is monotone, the exact integral (log Z, the thermodynamic identity) lies
    between the left and right Riemann sums, and the trapezoid, their mean, is
    within half their difference, 0.5 sum dalpha dE, of it.  On this data
    and schedule that is about 0.13 below log Z, inside the bound."""
    y = _k1_data()
    prior = oracle.gmm_prior(y)
    lp, ll, mu, u = _k1_tempered_grid(y, prior)
    exact = _logz_k1_quadrature(y, prior)
    llc = -0.5 * len(y) * LN2PI

    def e_loglik(a):
        f = lp + a * ll
        w = np.exp(f - f.max())
        return float((w * ll).sum() / w.sum())

    diffs = []
    for s in range(8):
        rec = oracle.gmm_run(16384, y, K=1, sweeps=3, seed=300 + s, resample_seed=400 + s,
                             threads=8, want_state=False)[3]
        a = rec["alpha"]
        el = np.array([e_loglik(x) for x in a])
        trap = float(np.sum(0.5 * np.diff(a) * (el[1:] + el[:-1]))) + llc
        diffs.append(rec["path_log_z"][-1] - trap)
        assert (np.diff(el) >= -1e-9).all()
        bound = 0.5 * float(np.sum(np.diff(a) * np.diff(el)))
        assert abs(trap - exact) <= bound + 1e-3, (trap - exact, bound)
    d = np.array(diffs)
    se = d.std(ddof=1) / np.sqrt(d.size)
    assert abs(d.mean()) < 4 * np.hypot(se, 2e-3), (d, se)
