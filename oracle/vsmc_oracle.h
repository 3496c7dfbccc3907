/*
 * vsmc_oracle.h -- CPU restatement of vSMC's per-iteration particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in vsmc_amd/ links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line in the
 * vSMC tree) in plain C, FP64, sequential where vSMC is sequential, with the
 * move loop parallelised by OpenMP the way MoveOMP does
 * (include/vsmc/smp/backend_omp.hpp:104-117).  Build flags must keep
 * -ffp-contract=off so that the uniform grids u0 + n*delta round as vSMC's
 * x86-64 build does.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - engine blocks / streams / discard  : golden vectors written by the
 *     compiled vSMC headers (tests/golden/ref_vectors.json) + Random123 KATs
 *   - generate_canonical, U01Sequence*    : golden vectors from the compiled
 *     headers (libstdc++ 11 uniform_real_distribution)
 *   - inversion / cfrp_trans / 6 schemes  : golden vectors (replication,
 *     copy_from, engine state after) + brute-force invariants
 *   - WeightSet normalisation / ESS       : golden vectors (bit-exact)
 *   - fold_prefix                         : definition (left fold) + exact
 *     rational check on tiny inputs
 *   - CV particle filter run              : golden run driven through
 *     vsmc::Particle<T>, WeightSet and Resample<Systematic>
 *   - GMM tempered SMC                    : the tempering search (CESS and
 *     alpha_ess bisection) equals the compiled move_alpha.hpp and WeightSet
 *     (golden fixtures and the live reference harness); log-likelihood by
 *     closed form; log evidence by quadrature for K = 1.  Initial draws and
 *     MH normals come from other engines than the reference's per-particle
 *     std distributions, so the whole run is pinned statistically only.
 *   - Kalman filter                       : closed form, checked against a
 *     brute-force Gaussian marginal on tiny T
 */
#ifndef VSMC_ORACLE_H
#define VSMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- counter-based RNG blocks ---------------- */
void orc_philox4x32_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_threefry4x64_block(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);
void orc_threefry4x32_block(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4]);

/* ---------------- engines (ctr, key, buffer, index) ---------------- */
typedef struct { uint64_t ctr[4]; uint64_t key[4]; uint64_t buf[4]; uint32_t index; } orc_tf64_engine;
typedef struct { uint32_t ctr[4]; uint32_t key[4]; uint32_t buf[4]; uint32_t index; } orc_tf32_engine;
typedef struct { uint32_t ctr[4]; uint32_t key[2]; uint32_t buf[4]; uint32_t index; } orc_ph32_engine;

void orc_tf64_seed(orc_tf64_engine *e, uint64_t s);
uint64_t orc_tf64_next(orc_tf64_engine *e);
void orc_tf64_discard(orc_tf64_engine *e, uint64_t n);
void orc_tf32_seed(orc_tf32_engine *e, uint32_t s);
uint32_t orc_tf32_next(orc_tf32_engine *e);
void orc_ph32_seed(orc_ph32_engine *e, uint32_t s);
uint32_t orc_ph32_next(orc_ph32_engine *e);
void orc_ph32_discard(orc_ph32_engine *e, uint64_t n);

/* engine-stream helpers for tests: n successive draws */
void orc_tf64_stream(orc_tf64_engine *e, size_t n, uint64_t *out);
void orc_ph32_stream(orc_ph32_engine *e, size_t n, uint32_t *out);
void orc_tf32_stream(orc_tf32_engine *e, size_t n, uint32_t *out);

/* ---------------- u01 ---------------- */
float orc_u01_open_closed_32_24(uint32_t u);
double orc_u01_open_closed_32_53(uint32_t u);
double orc_u01_closed_open_64_53(uint64_t u);
/* std::uniform_real_distribution<double>(0,1) on a 64-bit / 32-bit engine */
double orc_canonical_tf64(orc_tf64_engine *e);
double orc_canonical_ph32(orc_ph32_engine *e);

/* ---------------- resampling ---------------- */
enum {
    ORC_MULTINOMIAL = 0,
    ORC_RESIDUAL = 1,
    ORC_STRATIFIED = 2,
    ORC_SYSTEMATIC = 3,
    ORC_RESIDUAL_STRATIFIED = 4,
    ORC_RESIDUAL_SYSTEMATIC = 5
};

/* left fold a_i = fl(a_{i-1} + x_i), a_{-1} = 0 */
void orc_fold_prefix(size_t n, const double *x, double *out);
double orc_asum(size_t n, const double *x);
double orc_dot(size_t n, const double *x, const double *y);

/* inversion over an explicit uniform array; returns the number of uniforms
 * examined (the lazy-sequence draw count) */
size_t orc_inversion_u(size_t M, size_t N, const double *w, const double *u, uint64_t *rep);
void orc_cfrp_trans(size_t M, size_t N, const uint64_t *rep, uint64_t *copy_from);

/* Resample<Scheme>::operator()(M, N, rng, w, rep); returns draws made */
size_t orc_resample(int scheme, size_t M, size_t N, orc_tf64_engine *eng,
                    const double *w, uint64_t *rep);
/* the uniform sequence a scheme would produce, fully drawn (N values) */
void orc_u01seq(int kind /*0 sorted,1 stratified,2 systematic*/, size_t N,
                orc_tf64_engine *eng, double *u);
/* residual split: integral parts, scaled residuals (rr), and NN = N - R */
size_t orc_residual_split(size_t M, size_t N, const double *w, uint64_t *integral, double *rr);

/* ---------------- WeightSet ---------------- */
/* set_log_weight: lw normalised in place (max subtracted), w out, ESS returned */
double orc_weightset_set_log_weight(size_t N, double *lw, double *w);
double orc_weightset_add_log_weight(size_t N, double *lw, const double *inc, double *w);

/* ---------------- CV particle filter (config 0/1 model) ---------------- */
typedef struct {
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel;
    double obs_scale, nu, llh_const; /* llh_const = per-coordinate log normaliser */
} orc_cv_param;

typedef struct {
    double ess, log_z, mean_x, mean_y;
    int resampled;
    uint64_t draws;
} orc_step_record;

/* one FP64 Box-Muller block: 4 normals from philox(ctr={t+1,id_lo,id_hi,0}, key=seed) */
void orc_cv_normals(uint64_t seed, uint64_t id, uint64_t t, double z[4]);
double orc_cv_llh(const orc_cv_param *p, double x, double y, double ox, double oy);
void orc_cv_llh_many(const orc_cv_param *p, size_t n, const double *st, double ox, double oy,
                     double *inc);
void orc_cv_move_many(const orc_cv_param *p, uint64_t seed, size_t t, size_t n,
                      const double *st_in, double ox, double oy, double *st_out, double *inc);

/* Full filter run, FP64, vSMC step order (core/sampler.hpp:773-834).
 * state: N x 4 row-major (x, y, vx, vy) out; lw: N out; rec: T records.
 * obs: T x 2.  threads <= 0 -> OpenMP default. */
int orc_cv_run(const orc_cv_param *p, size_t N, size_t T, const double *obs,
               uint64_t seed, uint64_t resample_seed, int scheme, double threshold,
               double *state, double *lw, orc_step_record *rec, int threads);

/* timing-oriented variant of orc_cv_run: same arithmetic, returns per-phase
 * seconds in phase[4] = {move, normalise, resample, copy} */
/* backends of the timed CPU path: vSMC's OpenMP and std::thread SMP backends */
#define ORC_BACKEND_OMP 0
#define ORC_BACKEND_THREADS 1
#define ORC_MAX_THREADS 1024
int orc_cv_run_timed(const orc_cv_param *p, size_t N, size_t T, const double *obs,
                     uint64_t seed, uint64_t resample_seed, int scheme, double threshold,
                     orc_step_record *rec, int threads, int backend, double phase[4]);

/* ---------------- GMM tempered SMC (gmm_oracle.c) ---------------- */
#define ORC_GMM_MAX_K 6
typedef struct {
    int K;               /* mixture components, 1 .. ORC_GMM_MAX_K */
    uint32_t sweeps;     /* MH sweeps per iteration (each: mu, lambda, weight) */
    const double *data;
    size_t ndata;
    double threshold;    /* resample when ESS < threshold * N */
    double cess_drop;    /* next alpha keeps CESS = 1 - cess_drop */
    int scheme;          /* resampling scheme */
    uint64_t seed, resample_seed;
    size_t max_steps;
} orc_gmm_config;

typedef struct {
    double alpha, alpha_inc, ess, log_z, path_integrand, path_log_z;
    int32_t resampled;
    uint32_t bisections;
    uint64_t accept[3];
} orc_gmm_record;

/* prior = {mu0, sd0, shape0, scale0} from the data range */
void orc_gmm_prior(const double *data, size_t n, double prior[4]);
double orc_gmm_log_prior(int K, const double *mu, const double *lambda, const double prior[4]);
double orc_gmm_log_likelihood(int K, const double *mu, const double *lambda, const double *weight,
                              const double *data, size_t n);
/* p = {mu[K], lambda[K], weight[K]} */
void orc_gmm_init_particle(int K, uint64_t seed, uint64_t id, const double prior[4], double *p);
int orc_gmm_mh_move(int K, int move, uint64_t seed, uint64_t id, uint32_t iter, uint32_t sweep,
                    double alpha, const double sd[3], const double prior[4], const double *data,
                    size_t n, double *p, double *llh, double *lp);
/* Whole run (initialize, then iterate until alpha = 1 or max_steps).
 * params_out: N x 3K, llh_out, lw_out: N (any may be NULL); rec: max_steps + 1
 * rows (row 0 = initialize); *rows = rows written. */
double orc_gmm_cess(size_t N, const double *W, const double *llh, double a_diff);
uint32_t orc_gmm_alpha_search(size_t N, const double *W, const double *llh, double alpha,
                              double drop, double *alpha_out, double *alpha_inc_out);
int orc_gmm_run(const orc_gmm_config *c, size_t N, double *params_out, double *llh_out,
                double *lw_out, orc_gmm_record *rec, size_t *rows, int threads);

/* ---- OpenCL-C RNG layer and the OpenCL pf example (user-kernel convention) ---- */
typedef struct { uint32_t key[4], ctr[4], rnd[4]; unsigned remain; } orc_cburng4x32;
typedef struct { double u1, u2; int saved; } orc_normal01_53;
void orc_cburng4x32_init(orc_cburng4x32 *r, uint32_t seed);
uint32_t orc_cburng4x32_rand(orc_cburng4x32 *r);
void orc_normal01_4x32_53_init(orc_normal01_53 *n, orc_cburng4x32 *r);
double orc_normal01_4x32_53_rand(orc_normal01_53 *n, orc_cburng4x32 *r);
void orc_cburng4x32_stream(uint32_t seed, size_t n, uint32_t *out);
void orc_normal01_4x32_53_stream(uint32_t seed, size_t n, double *out);
int orc_clpf_run(size_t N, size_t T, const double *obs, uint32_t seed, uint64_t resample_seed,
                 int scheme, double threshold, double *state_out, orc_step_record *rec);

#ifdef __cplusplus
}
#endif
#endif
