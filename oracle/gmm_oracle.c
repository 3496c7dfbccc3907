/* gmm_oracle.c -- CPU restatement (FP64) of vSMC's Gaussian-mixture SMC
 * sampler with adaptive (CESS) tempering: example/gmm/src/gmm_smc.cpp with
 * the "cess" schedule of example/include/smc_do.hpp.
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker of the HIP path; the product
 * never links it.
 *
 * What it follows, line by line:
 *   prior constants from the data range       example/gmm/include/gmm_state.hpp:120-160
 *   log_prior / log_likelihood / log_target   example/gmm/include/gmm_state.hpp:73-116
 *   init (normal mu, gamma lambda, Dirichlet) example/gmm/include/gmm_init.hpp:40-70
 *   alpha update (clamp, alpha_inc)           example/gmm/include/gmm_state.hpp:163-172
 *   CESS bisection (cess01, target 1 - drop)  example/include/move_alpha.hpp:203-231, 242-279
 *   CESS = (sum W b)^2 / sum W b^2            include/vsmc/core/weight_set.hpp:411-431
 *   proposal scales alpha2sd                  example/gmm/include/gmm_proposal.hpp:71-80
 *   reweight, log-evidence increment          example/include/move_smc.hpp:57-77
 *   MH moves mu, lambda (lognormal), weight   example/gmm/include/gmm_move.hpp:35-121
 *     with their Jacobians                    example/gmm/include/gmm_param.hpp:202-248
 *   accept iff p >= log u                     example/gmm/include/gmm_param.hpp:250-324
 *   iteration order move, resample, mcmc      include/vsmc/core/sampler.hpp:773-834
 *   path sampling (trapezoid on alpha)        include/vsmc/core/path.hpp:229-239,
 *                                             example/include/smc_path.hpp:40-44
 *   output log Z = zconst + log_likelihood_const   example/include/smc_do.hpp:75-79
 *
 * Random numbers.  The reference draws from per-particle C++ engines through
 * std::normal/gamma/lognormal/uniform distributions; this restatement (and
 * the HIP path it checks) draws from Philox4x32-10 blocks keyed by the seed
 * with counter {iteration, id_lo, id_hi, block}:
 *   init (iteration 0): word stream of blocks 0..6 --
 *     words 0..7   Box-Muller pairs -> mu_d = mu0 + sd0 * z_d
 *     words 8..19  lambda_d = scale0 * (-log U_{8+2d} - log U_{9+2d})  (Gamma(2, scale0):
 *                  shape0 = 2 is fixed by the reference, a sum of two exponentials)
 *     words 20..25 g_d = -log U_{20+d} (Gamma(1, 1)); w_d = g_d / sum g
 *   MH (iteration t >= 1), sweep s, move b (0 mu, 1 lambda, 2 weight):
 *     blocks 8 + 2 (3 s + b) and the next one: words 0..5 -> three Box-Muller
 *     pairs, word 7 -> U for log u.
 * U = u01_open_closed_32_53 (include/vsmc/rng/u01.h:139-141), Box-Muller as
 * include/vsmc/rng/normal01.h:118-139.  The distributions are the
 * reference's; the variate streams are this build's.
 *
 * The log-evidence increment log sum_i W_i exp(inc_i) is evaluated with a
 * max shift (the reference evaluates it unshifted; the two agree up to
 * rounding while exp(inc_i) does not underflow).  The CESS bisection keeps
 * the reference's unshifted exp(a_diff * llh_i), including its behaviour
 * when every b_i underflows (0/0 is NaN, which is not < target, so the
 * search moves up).
 */
#include "vsmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GMM_LN_2PI 1.837877066409345483560659472811235279723

void orc_gmm_prior(const double *data, size_t n, double prior[4])
{
    double xmax = data[0], xmin = data[0];
    for (size_t i = 0; i < n; ++i) {
        if (data[i] > xmax) xmax = data[i];
        if (data[i] < xmin) xmin = data[i];
    }
    const double kappa = 1 / ((xmax - xmin) * (xmax - xmin));
    prior[0] = 0.5 * (xmin + xmax); /* mu0 */
    prior[1] = 1 / sqrt(kappa);     /* sd0 */
    prior[2] = 2;                   /* shape0 */
    prior[3] = 50 * kappa;          /* scale0 */
}

static void gmm_block(uint64_t seed, uint64_t id, uint32_t iter, uint32_t blk, uint32_t w[4])
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {iter, (uint32_t)id, (uint32_t)(id >> 32), blk};
    orc_philox4x32_block(ctr, key, w);
}

static void box_muller(uint32_t a, uint32_t b, double *z0, double *z1)
{
    const double two_pi = 6.2831853071795865;
    const double u1 = orc_u01_open_closed_32_53(a);
    const double u2 = orc_u01_open_closed_32_53(b);
    const double r = sqrt(-2 * log(u1));
    *z0 = r * cos(two_pi * u2);
    *z1 = r * sin(two_pi * u2);
}

/* gmm_state.hpp:81-95 */
double orc_gmm_log_prior(int K, const double *mu, const double *lambda, const double prior[4])
{
    const double mu0 = prior[0], sd0 = prior[1], shape0 = prior[2], scale0 = prior[3];
    double lp = 0;
    for (int d = 0; d < K; ++d) {
        const double resid = mu[d] - mu0;
        lp += -0.5 * (resid * resid) / (sd0 * sd0);
        lp += (shape0 - 1) * log(lambda[d]) - lambda[d] / scale0;
    }
    return lp;
}

/* gmm_state.hpp:97-116 */
double orc_gmm_log_likelihood(int K, const double *mu, const double *lambda, const double *weight,
                              const double *data, size_t n)
{
    double ll_lambda[ORC_GMM_MAX_K];
    for (int d = 0; d < K; ++d) ll_lambda[d] = log(lambda[d]);
    double ll = 0;
    for (size_t i = 0; i < n; ++i) {
        double lli = 0;
        for (int d = 0; d < K; ++d) {
            const double resid = data[i] - mu[d];
            lli += weight[d] * exp(0.5 * ll_lambda[d] - 0.5 * lambda[d] * resid * resid);
        }
        ll += log(lli + 1e-13);
    }
    return ll;
}

/* gmm_param.hpp:221-248 */
static double logit_part(int K, const double *w)
{
    double sum_lw = 1, sum_llw = 0;
    for (int d = 0; d < K - 1; ++d) {
        const double r = w[d] / w[K - 1];
        sum_lw += r;
        sum_llw += log(r);
    }
    return sum_llw - K * log(sum_lw);
}

void orc_gmm_init_particle(int K, uint64_t seed, uint64_t id, const double prior[4], double *p)
{
    uint32_t w[28];
    for (uint32_t b = 0; b < 7; ++b) gmm_block(seed, id, 0, b, w + 4 * b);
    double *mu = p, *lambda = p + K, *weight = p + 2 * K;
    double z[8];
    for (int k = 0; k < 4; ++k) box_muller(w[2 * k], w[2 * k + 1], z + 2 * k, z + 2 * k + 1);
    double sum = 0;
    for (int d = 0; d < K; ++d) {
        mu[d] = prior[0] + prior[1] * z[d];
        lambda[d] = prior[3] * (-log(orc_u01_open_closed_32_53(w[8 + 2 * d])) -
                                log(orc_u01_open_closed_32_53(w[9 + 2 * d])));
        weight[d] = -log(orc_u01_open_closed_32_53(w[20 + d]));
        sum += weight[d];
    }
    for (int d = 0; d < K; ++d) weight[d] /= sum;
}

/* One MH move (gmm_move.hpp:35-121) of one particle: p = 3K parameters,
 * llh / lp its current likelihood and prior.  Returns 1 if accepted. */
int orc_gmm_mh_move(int K, int move, uint64_t seed, uint64_t id, uint32_t iter, uint32_t sweep,
                    double alpha, const double sd[3], const double prior[4], const double *data,
                    size_t n, double *p, double *llh, double *lp)
{
    uint32_t w[8];
    const uint32_t blk = 8 + 2 * (3 * sweep + (uint32_t)move);
    gmm_block(seed, id, iter, blk, w);
    gmm_block(seed, id, iter, blk + 1, w + 4);
    double z[6];
    for (int k = 0; k < 3; ++k) box_muller(w[2 * k], w[2 * k + 1], z + 2 * k, z + 2 * k + 1);
    const double log_u = log(orc_u01_open_closed_32_53(w[7]));
    double old[3 * ORC_GMM_MAX_K];
    memcpy(old, p, sizeof(double) * 3 * K);
    double *mu = p, *lambda = p + K, *weight = p + 2 * K;
    double jac = 0;
    if (move == 0) {
        for (int d = 0; d < K; ++d) mu[d] += sd[0] * z[d];
    } else if (move == 1) {
        for (int d = 0; d < K; ++d) lambda[d] *= exp(sd[1] * z[d]);
        for (int d = 0; d < K; ++d) jac += log(lambda[d]) - log(old[K + d]);
    } else {
        double sum_weight = 1;
        for (int d = 0; d < K - 1; ++d) {
            weight[d] = log(weight[d] / weight[K - 1]);
            weight[d] += sd[2] * z[d];
            weight[d] = exp(weight[d]);
            sum_weight += weight[d];
        }
        weight[K - 1] = 1;
        for (int d = 0; d < K; ++d) weight[d] /= sum_weight;
        jac = logit_part(K, weight) - logit_part(K, old + 2 * K);
    }
    const double lp_new = orc_gmm_log_prior(K, mu, lambda, prior);
    const double ll_new = orc_gmm_log_likelihood(K, mu, lambda, weight, data, n);
    const double pr = (lp_new + alpha * ll_new) - (*lp + alpha * *llh) + jac;
    if (pr < log_u) {
        memcpy(p, old, sizeof(double) * 3 * K);
        return 0;
    }
    *llh = ll_new;
    *lp = lp_new;
    return 1;
}

/* gmm_proposal.hpp:71-80 */
static void alpha2sd(double alpha, double sd[3])
{
    alpha = alpha < 0.05 ? 0.05 : alpha;
    sd[0] = 0.15 / alpha;
    sd[1] = (1 + sqrt(1.0 / alpha)) * 0.15;
    sd[2] = (1 + sqrt(1.0 / alpha)) * 0.2;
}

/* cess01 (move_alpha.hpp:203-211) -> WeightSet::compute_cess with use_log
 * false (include/vsmc/core/weight_set.hpp:411-431): the multipliers are
 * b_i = exp(a_diff * llh_i) unshifted (move_alpha.hpp:256-262), and
 * CESS = (sum_i W_i b_i)^2 / sum_i W_i b_i^2 accumulated in index order.
 *
 * ORC_PLANT_CESS (never set by oracle/Makefile) plants a plausible mistake,
 * for the tests that show the golden vectors catch it:
 *   1  the ESS of the reweighted weights instead of the CESS (ess01), with the
 *      target (1 - drop) * ESS of the current weights (move_alpha.hpp:213-222)
 *   2  use_log semantics: the multipliers exponentiated a second time
 *   3  the search started at alpha + 0.1 instead of alpha + 0.05 */
double orc_gmm_cess(size_t N, const double *W, const double *llh, double a_diff)
{
    double above = 0, below = 0;
    for (size_t i = 0; i < N; ++i) {
        double b = exp(a_diff * llh[i]);
#if defined(ORC_PLANT_CESS) && ORC_PLANT_CESS == 2
        b = exp(b);
#endif
        const double wb = W[i] * b;
        above += wb;
#if defined(ORC_PLANT_CESS) && ORC_PLANT_CESS == 1
        below += wb * wb;
#else
        below += wb * b;
#endif
    }
    return above * above / below;
}

/* alpha_ess<T, cess01<T>>::alpha_iter (move_alpha.hpp:242-279): bisection on
 * [alpha, 1.05] from alpha + 0.05 until the interval is within
 * e = max(1e-10, 1e-3 drop) or the lower end passes 1; then the clamp to
 * [0, 1] and alpha_inc of gmm_state::alpha (gmm_state.hpp:163-172).  W are
 * the current normalised weights.  Returns the number of CESS evaluations. */
uint32_t orc_gmm_alpha_search(size_t N, const double *W, const double *llh, double alpha,
                              double drop, double *alpha_out, double *alpha_inc_out)
{
    const double e = 1e-10 > 1e-3 * drop ? 1e-10 : 1e-3 * drop;
    double target = 1 - drop;
    double l = alpha, u = 1.05, a = l + 0.05;
#if defined(ORC_PLANT_CESS) && ORC_PLANT_CESS == 3
    a = l + 0.1;
#endif
#if defined(ORC_PLANT_CESS) && ORC_PLANT_CESS == 1
    {
        double ss = 0;
        for (size_t i = 0; i < N; ++i) ss += W[i] * W[i];
        target = (1 - drop) / ss;
    }
#endif
    const double a_old = l;
    uint32_t iters = 0;
    while (fabs(u - l) > e && l <= 1) {
        const double cess = orc_gmm_cess(N, W, llh, a - a_old);
        ++iters;
        if (cess < target) {
            u = a;
            a = 0.5 * (a + l);
        } else {
            l = a;
            a = 0.5 * (a + u);
        }
    }
    a = a < 1 ? a : 1;
    a = a > 0 ? a : 0;
    *alpha_inc_out = a > 0 ? a - alpha : 0;
    *alpha_out = a;
    return iters;
}

int orc_gmm_run(const orc_gmm_config *c, size_t N, double *params_out, double *llh_out,
                double *lw_out, orc_gmm_record *rec, size_t *rows, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int K = c->K;
    if (K < 1 || K > ORC_GMM_MAX_K || N == 0) return -2;
    const size_t P = 3 * (size_t)K;
    double prior[4];
    orc_gmm_prior(c->data, c->ndata, prior);
    const double llc = -0.5 * (double)c->ndata * GMM_LN_2PI;
    double *p = (double *)malloc(sizeof(double) * P * N);
    double *p2 = (double *)malloc(sizeof(double) * P * N);
    double *llh = (double *)malloc(sizeof(double) * N);
    double *llh2 = (double *)malloc(sizeof(double) * N);
    double *lp = (double *)malloc(sizeof(double) * N);
    double *lp2 = (double *)malloc(sizeof(double) * N);
    double *lw = (double *)malloc(sizeof(double) * N);
    double *W = (double *)malloc(sizeof(double) * N);
    double *inc = (double *)malloc(sizeof(double) * N);
    uint64_t *rep = (uint64_t *)malloc(sizeof(uint64_t) * N);
    uint64_t *cf = (uint64_t *)malloc(sizeof(uint64_t) * N);
    uint64_t *acc = (uint64_t *)calloc(3 * N, sizeof(uint64_t));
    int rc = 0;
    if (!p || !p2 || !llh || !llh2 || !lp || !lp2 || !lw || !W || !inc || !rep || !cf || !acc) {
        rc = -1;
        goto out;
    }
    orc_tf64_engine reng;
    orc_tf64_seed(&reng, c->resample_seed);

    /* initialize: equal weights, alpha = 0, draws from the prior, log_target */
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ++ii) {
        const size_t i = (size_t)ii;
        double *q = p + P * i;
        orc_gmm_init_particle(K, c->seed, i, prior, q);
        lp[i] = orc_gmm_log_prior(K, q, q + K, prior);
        llh[i] = orc_gmm_log_likelihood(K, q, q + K, q + 2 * K, c->data, c->ndata);
    }
    for (size_t i = 0; i < N; ++i) { lw[i] = 0; W[i] = 1.0 / (double)N; }
    double alpha = 0, zconst = 0, path = 0;
    double integrand = 0;
    for (size_t i = 0; i < N; ++i) integrand += W[i] * llh[i];
    {
        orc_gmm_record r;
        memset(&r, 0, sizeof(r));
        r.ess = (double)N;
        r.log_z = llc;
        r.path_integrand = integrand;
        r.path_log_z = llc;
        rec[0] = r;
    }
    size_t t = 0;
    while (alpha < 1 && t < c->max_steps) {
        ++t;
        /* move_smc: alpha_iter (cess bisection) */
        if (alpha <= 0) zconst = 0;
        double a, a_inc;
        const uint32_t iters = orc_gmm_alpha_search(N, W, llh, alpha, c->cess_drop, &a, &a_inc);
        alpha = a;
        double sd[3];
        alpha2sd(alpha, sd);
        /* reweight: log sum_i W_i exp(inc_i), inc = llh * alpha_inc */
        double imax = -INFINITY;
        for (size_t i = 0; i < N; ++i) {
            inc[i] = llh[i] * a_inc;
            if (imax < inc[i]) imax = inc[i];
        }
        double sum = 0;
        for (size_t i = 0; i < N; ++i) sum += W[i] * exp(inc[i] - imax);
        zconst += log(sum) + imax;
        const double ess = orc_weightset_add_log_weight(N, lw, inc, W);
        /* resample (stratified in the reference sampler) */
        const int resampled = ess < c->threshold * (double)N;
        if (resampled) {
            orc_resample(c->scheme, N, N, &reng, W, rep);
            orc_cfrp_trans(N, N, rep, cf);
            for (size_t to = 0; to < N; ++to) {
                const size_t from = cf[to];
                memcpy(p2 + P * to, p + P * from, sizeof(double) * P);
                llh2[to] = llh[from];
                lp2[to] = lp[from];
            }
            double *tp = p; p = p2; p2 = tp;
            tp = llh; llh = llh2; llh2 = tp;
            tp = lp; lp = lp2; lp2 = tp;
            for (size_t i = 0; i < N; ++i) { W[i] = 1.0 / (double)N; lw[i] = 0; }
        }
        /* mcmc: sweeps x (mu, lambda, weight), each particle independent */
#pragma omp parallel for schedule(static)
        for (long long ii = 0; ii < (long long)N; ++ii) {
            const size_t i = (size_t)ii;
            for (uint32_t s = 0; s < c->sweeps; ++s)
                for (int b = 0; b < 3; ++b)
                    acc[3 * i + b] += (uint64_t)orc_gmm_mh_move(
                        K, b, c->seed, i, (uint32_t)t, s, alpha, sd, prior, c->data, c->ndata,
                        p + P * i, llh + i, lp + i);
        }
        /* path sampling at MonitorMCMC */
        double integ = 0;
        for (size_t i = 0; i < N; ++i) integ += W[i] * llh[i];
        path += 0.5 * (alpha - rec[t - 1].alpha) * (integ + integrand);
        integrand = integ;
        orc_gmm_record r;
        memset(&r, 0, sizeof(r));
        r.alpha = alpha;
        r.alpha_inc = a_inc;
        r.ess = ess;
        r.log_z = zconst + llc;
        r.path_integrand = integ;
        r.path_log_z = path + llc;
        r.resampled = resampled;
        r.bisections = iters;
        for (int b = 0; b < 3; ++b) {
            uint64_t s = 0;
            for (size_t i = 0; i < N; ++i) { s += acc[3 * i + b]; acc[3 * i + b] = 0; }
            r.accept[b] = s;
        }
        rec[t] = r;
    }
    *rows = t + 1;
    if (params_out) memcpy(params_out, p, sizeof(double) * P * N);
    if (llh_out) memcpy(llh_out, llh, sizeof(double) * N);
    if (lw_out) memcpy(lw_out, lw, sizeof(double) * N);
out:
    free(p); free(p2); free(llh); free(llh2); free(lp); free(lp2); free(lw); free(W); free(inc);
    free(rep); free(cf); free(acc);
    return rc;
}
