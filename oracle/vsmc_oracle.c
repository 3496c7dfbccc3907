/*
 * vsmc_oracle.c -- CPU restatement of vSMC's per-iteration particle step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library.  The product path
 * (vsmc_amd/) never links, loads or calls it.
 *
 * Compile with -ffp-contract=off: the uniform grids (u0 + n*delta) and the
 * running sums must round exactly as vSMC's x86-64 build rounds them.
 * Reference paths are relative to the vSMC source tree.
 */
#include "vsmc_oracle.h"

#include <pthread.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * Philox4x32-10 -- include/vsmc/rng/philox.hpp
 *   Weyl constants 0x9E3779B9 / 0xBB67AE85            (philox.hpp:77-80)
 *   multipliers M0 = 0xD2511F53, M1 = 0xCD9E8D57       (philox.hpp:92-95)
 *   round: (hi0,lo1) = M1*s2, (hi2,lo3) = M0*s0,
 *          s = {hi0^k0^s1, lo1, hi2^k1^s3, lo3}        (philox.hpp:245-265)
 *   10 rounds, key bumped before rounds 2..10          (philox.hpp:504-523)
 * ====================================================================== */
void orc_philox4x32_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t s0 = ctr[0], s1 = ctr[1], s2 = ctr[2], s3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 1; r <= 10; ++r) {
        if (r > 1) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p1 = (uint64_t)0xCD9E8D57u * s2;
        uint64_t p0 = (uint64_t)0xD2511F53u * s0;
        uint32_t hi0 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t hi2 = (uint32_t)(p0 >> 32), lo3 = (uint32_t)p0;
        uint32_t n0 = hi0 ^ k0 ^ s1;
        uint32_t n2 = hi2 ^ k1 ^ s3;
        s0 = n0; s1 = lo1; s2 = n2; s3 = lo3;
    }
    out[0] = s0; out[1] = s1; out[2] = s2; out[3] = s3;
}

/* ======================================================================
 * Threefry4x64-20 / Threefry4x32-20 -- include/vsmc/rng/threefry.hpp
 *   KS parity constant                                 (threefry.hpp:71-73)
 *   rotation constants R[(N-1)%8][0..1]                (threefry.hpp:87-130)
 *   mix: N odd -> pairs (0,1),(2,3); N even -> (0,3),(2,1)
 *                                                      (threefry.hpp:162-185)
 *   key injection every 4 rounds, inc = N/4            (threefry.hpp:213-233)
 *   par = {k0..k3, KS ^ k0 ^ .. ^ k3}                   (threefry.hpp:491-506)
 * ====================================================================== */
static const unsigned orc_tf64_rot[8][2] = {
    {14, 16}, {52, 57}, {23, 40}, {5, 37}, {25, 33}, {46, 12}, {58, 22}, {32, 32}};
static const unsigned orc_tf32_rot[8][2] = {
    {10, 26}, {11, 21}, {13, 27}, {23, 5}, {6, 20}, {17, 11}, {25, 10}, {18, 20}};

static inline uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }
static inline uint32_t rotl32(uint32_t x, unsigned r) { return (x << r) | (x >> (32 - r)); }

void orc_threefry4x64_block(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4])
{
    uint64_t par[5];
    par[4] = UINT64_C(0x1BD11BDAA9FC1A22);
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint64_t s[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    for (int j = 0; j < 4; ++j) s[j] += par[j];
    for (unsigned n = 1; n <= 20; ++n) {
        unsigned r = (n - 1) % 8;
        unsigned i0 = (n % 2) ? 1 : 3, i2 = (n % 2) ? 3 : 1;
        s[0] += s[i0];
        s[i0] = rotl64(s[i0], orc_tf64_rot[r][0]);
        s[i0] ^= s[0];
        s[2] += s[i2];
        s[i2] = rotl64(s[i2], orc_tf64_rot[r][1]);
        s[i2] ^= s[2];
        if (n % 4 == 0) {
            unsigned inc = n / 4;
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    for (int j = 0; j < 4; ++j) out[j] = s[j];
}

void orc_threefry4x32_block(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4])
{
    uint32_t par[5];
    par[4] = 0x1BD11BDAu;
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint32_t s[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    for (int j = 0; j < 4; ++j) s[j] += par[j];
    for (unsigned n = 1; n <= 20; ++n) {
        unsigned r = (n - 1) % 8;
        unsigned i0 = (n % 2) ? 1 : 3, i2 = (n % 2) ? 3 : 1;
        s[0] += s[i0];
        s[i0] = rotl32(s[i0], orc_tf32_rot[r][0]);
        s[i0] ^= s[0];
        s[2] += s[i2];
        s[i2] = rotl32(s[i2], orc_tf32_rot[r][1]);
        s[i2] ^= s[2];
        if (n % 4 == 0) {
            unsigned inc = n / 4;
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    for (int j = 0; j < 4; ++j) out[j] = s[j];
}

/* ======================================================================
 * Engines.  Counter ripple increment, word 0 least significant
 * (utility/counter.hpp:180-192); operator() increments the counter before
 * generating a block and returns buffer words 0..3 in order
 * (philox.hpp:386-395, threefry.hpp:355-364); seed(s) -> ctr = 0,
 * key = {s, 0, ...}, index = K (philox.hpp:344-350, threefry.hpp:302-310);
 * discard (philox.hpp:412-432, threefry.hpp:381-402).
 * ====================================================================== */
#define ORC_INC4(ctr)                                                   \
    do {                                                                \
        if (++(ctr)[0] == 0 && ++(ctr)[1] == 0 && ++(ctr)[2] == 0)      \
            ++(ctr)[3];                                                 \
    } while (0)

/* ctr += n as a 4-word integer with 64-bit words */
static void orc_add_ctr64(uint64_t ctr[4], uint64_t n)
{
    uint64_t old = ctr[0];
    ctr[0] += n;
    if (ctr[0] < old && ++ctr[1] == 0 && ++ctr[2] == 0) ++ctr[3];
}
static void orc_add_ctr32(uint32_t ctr[4], uint64_t n)
{
    uint64_t lo = (uint64_t)ctr[0] + (uint32_t)n;
    ctr[0] = (uint32_t)lo;
    uint64_t carry = (lo >> 32) + (n >> 32);
    for (int j = 1; j < 4 && carry; ++j) {
        uint64_t v = (uint64_t)ctr[j] + (uint32_t)carry;
        ctr[j] = (uint32_t)v;
        carry = (v >> 32) + (carry >> 32);
    }
}

void orc_tf64_seed(orc_tf64_engine *e, uint64_t s)
{
    memset(e, 0, sizeof(*e));
    e->key[0] = s;
    e->index = 4;
}
uint64_t orc_tf64_next(orc_tf64_engine *e)
{
    if (e->index == 4) {
        ORC_INC4(e->ctr);
        orc_threefry4x64_block(e->ctr, e->key, e->buf);
        e->index = 0;
    }
    return e->buf[e->index++];
}
void orc_tf64_discard(orc_tf64_engine *e, uint64_t n)
{
    if (e->index + n <= 4) { e->index += (uint32_t)n; return; }
    n -= 4 - e->index;
    if (n <= 4) { e->index = 4; orc_tf64_next(e); e->index = (uint32_t)n; return; }
    orc_add_ctr64(e->ctr, n / 4);
    e->index = 4;
    orc_tf64_next(e);
    e->index = (uint32_t)(n % 4);
}
void orc_tf32_seed(orc_tf32_engine *e, uint32_t s)
{
    memset(e, 0, sizeof(*e));
    e->key[0] = s;
    e->index = 4;
}
uint32_t orc_tf32_next(orc_tf32_engine *e)
{
    if (e->index == 4) {
        ORC_INC4(e->ctr);
        orc_threefry4x32_block(e->ctr, e->key, e->buf);
        e->index = 0;
    }
    return e->buf[e->index++];
}
void orc_ph32_seed(orc_ph32_engine *e, uint32_t s)
{
    memset(e, 0, sizeof(*e));
    e->key[0] = s;
    e->index = 4;
}
uint32_t orc_ph32_next(orc_ph32_engine *e)
{
    if (e->index == 4) {
        ORC_INC4(e->ctr);
        orc_philox4x32_block(e->ctr, e->key, e->buf);
        e->index = 0;
    }
    return e->buf[e->index++];
}
void orc_ph32_discard(orc_ph32_engine *e, uint64_t n)
{
    if (e->index + n <= 4) { e->index += (uint32_t)n; return; }
    n -= 4 - e->index;
    if (n <= 4) { e->index = 4; orc_ph32_next(e); e->index = (uint32_t)n; return; }
    orc_add_ctr32(e->ctr, n / 4);
    e->index = 4;
    orc_ph32_next(e);
    e->index = (uint32_t)(n % 4);
}
void orc_tf64_stream(orc_tf64_engine *e, size_t n, uint64_t *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_tf64_next(e);
}
void orc_ph32_stream(orc_ph32_engine *e, size_t n, uint32_t *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_ph32_next(e);
}
void orc_tf32_stream(orc_tf32_engine *e, size_t n, uint32_t *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_tf32_next(e);
}

/* ======================================================================
 * u01 conversions -- include/vsmc/rng/u01.h:103-167
 * ====================================================================== */
float orc_u01_open_closed_32_24(uint32_t u)
{
    /* (1 + (u >> 8)) * 2^-24  (u01.h:114-116) */
    return (1.0f + (float)(u >> 8)) * (1.0f / 16777216.0f);
}
double orc_u01_open_closed_32_53(uint32_t u)
{
    /* (1.0 + u) * 2^-32  (u01.h:139-141) */
    return (1.0 + (double)u) * (1.0 / 4294967296.0);
}
double orc_u01_closed_open_64_53(uint64_t u)
{
    /* (u >> 11) * 2^-53  (u01.h:159-161) */
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

/* std::uniform_real_distribution<double>(0,1)(eng) as libstdc++ 11 computes
 * it (cxx11/random.hpp:63 -> bits/random.tcc generate_canonical):
 *   64-bit engine: one draw, double(x) / 2^64;
 *   32-bit engine: two draws, (double(x0) + double(x1) * 2^32) / 2^64;
 *   a result >= 1 is replaced by nextafter(1, 0); then *(b-a) + a = *1 + 0. */
double orc_canonical_tf64(orc_tf64_engine *e)
{
    double sum = (double)orc_tf64_next(e) * 1.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r * 1.0 + 0.0;
}
double orc_canonical_ph32(orc_ph32_engine *e)
{
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum += (double)orc_ph32_next(e) * tmp;
        tmp *= 4294967296.0;
    }
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r * 1.0 + 0.0;
}

/* ======================================================================
 * BLAS-like folds -- include/vsmc/math/cblas.hpp:53-77 (sequential left folds)
 * ====================================================================== */
double orc_asum(size_t n, const double *x)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += fabs(x[i]);
    return s;
}
double orc_dot(size_t n, const double *x, const double *y)
{
    double s = 0;
    for (size_t i = 0; i < n; ++i) s = s + x[i] * y[i];
    return s;
}
void orc_fold_prefix(size_t n, const double *x, double *out)
{
    double a = 0;
    for (size_t i = 0; i < n; ++i) {
        a += x[i];
        out[i] = a;
    }
}

/* ======================================================================
 * Lazy uniform sequences -- include/vsmc/rng/u01.hpp:165-287.  Each draws
 * from the engine only when asked for an index it has not produced yet, so
 * the number of draws depends on how far inversion reads.
 * ====================================================================== */
typedef struct {
    int kind; /* 0 sorted, 1 stratified, 2 systematic */
    size_t N, n;
    double u, lmax, delta, u0;
    orc_tf64_engine *eng;
    size_t draws;
} orc_seq;

static void orc_seq_init(orc_seq *s, int kind, size_t N, orc_tf64_engine *eng)
{
    memset(s, 0, sizeof(*s));
    s->kind = kind;
    s->N = N;
    s->n = N;
    s->eng = eng;
    s->delta = 1.0 / (double)N;
    if (kind == 2) { /* U01SequenceSystematic ctor draws once (u01.hpp:258-263) */
        s->u0 = orc_canonical_tf64(eng) * s->delta;
        s->draws = 1;
    }
}

static double orc_seq_get(orc_seq *s, size_t n)
{
    if (n == s->n) return s->u;
    if (s->kind == 0) {
        /* lmax += log(1 - U) / (N - n); u = 1 - exp(lmax)  (u01.hpp:185-195) */
        s->lmax += log(1 - orc_canonical_tf64(s->eng)) / (double)(s->N - n);
        s->draws++;
        s->n = n;
        s->u = 1 - exp(s->lmax);
    } else if (s->kind == 1) {
        /* u = U * delta + n * delta  (u01.hpp:224-229) */
        s->n = n;
        s->u = orc_canonical_tf64(s->eng) * s->delta + (double)n * s->delta;
        s->draws++;
    } else {
        /* u = u0 + n * delta  (u01.hpp:270-274) */
        s->n = n;
        s->u = s->u0 + (double)n * s->delta;
    }
    return s->u;
}

void orc_u01seq(int kind, size_t N, orc_tf64_engine *eng, double *u)
{
    orc_seq s;
    orc_seq_init(&s, kind, N, eng);
    for (size_t n = 0; n < N; ++n) u[n] = orc_seq_get(&s, n);
}

/* internal::inversion -- include/vsmc/resample/common.hpp:57-84 */
static void orc_inversion_seq(size_t M, size_t N, const double *w, orc_seq *seq, uint64_t *rep)
{
    if (M == 0) return;
    if (M == 1) { rep[0] = N; return; }
    memset(rep, 0, sizeof(uint64_t) * M);
    if (N == 0) return;
    size_t n = 0;
    double accw = 0;
    for (size_t i = 0; i != M - 1; ++i) {
        accw += w[i];
        while (n != N && orc_seq_get(seq, n) <= accw) {
            ++rep[i];
            ++n;
        }
    }
    rep[M - 1] = N - n;
}

size_t orc_inversion_u(size_t M, size_t N, const double *w, const double *u, uint64_t *rep)
{
    if (M == 0) return 0;
    if (M == 1) { rep[0] = N; return 0; }
    memset(rep, 0, sizeof(uint64_t) * M);
    if (N == 0) return 0;
    size_t n = 0, seen = 0;
    double accw = 0;
    for (size_t i = 0; i != M - 1; ++i) {
        accw += w[i];
        while (n != N) {
            if (n + 1 > seen) seen = n + 1;
            if (!(u[n] <= accw)) break;
            ++rep[i];
            ++n;
        }
    }
    rep[M - 1] = N - n;
    return seen;
}

/* internal::cfrp_trans -- include/vsmc/resample/common.hpp:87-118 */
void orc_cfrp_trans(size_t M, size_t N, const uint64_t *rep, uint64_t *copy_from)
{
    if (M == N) {
        uint64_t time = 0;
        uint64_t from = 0;
        for (size_t to = 0; to != N; ++to) {
            if (rep[to] != 0) {
                copy_from[to] = to;
            } else {
                if (rep[from] < time + 2) {
                    time = 0;
                    do
                        ++from;
                    while (rep[from] < 2);
                }
                copy_from[to] = from;
                ++time;
            }
        }
    } else {
        size_t to = 0;
        for (size_t from = 0; from != M; ++from)
            for (uint64_t r = 0; r != rep[from]; ++r) copy_from[to++] = from;
    }
}

/* residual split -- include/vsmc/resample/residual.hpp:57-71:
 * r_i = modf(N * w_i, &i_i); r *= 1 / asum(r); NN = N - sum(i_i) */
size_t orc_residual_split(size_t M, size_t N, const double *w, uint64_t *integral, double *rr)
{
    uint64_t R = 0;
    for (size_t i = 0; i < M; ++i) {
        double ip;
        rr[i] = modf((double)N * w[i], &ip);
        integral[i] = (uint64_t)ip;
        R += integral[i];
    }
    double coeff = 1 / orc_asum(M, rr);
    for (size_t i = 0; i < M; ++i) rr[i] *= coeff;
    return N - (size_t)R;
}

/* Resample<Scheme>::operator() -- include/vsmc/resample/{multinomial,
 * residual,stratified,systematic,residual_stratified,residual_systematic}.hpp:53-77 */
size_t orc_resample(int scheme, size_t M, size_t N, orc_tf64_engine *eng,
                    const double *w, uint64_t *rep)
{
    orc_seq seq;
    if (scheme == ORC_MULTINOMIAL || scheme == ORC_STRATIFIED || scheme == ORC_SYSTEMATIC) {
        int kind = scheme == ORC_MULTINOMIAL ? 0 : (scheme == ORC_STRATIFIED ? 1 : 2);
        orc_seq_init(&seq, kind, N, eng);
        orc_inversion_seq(M, N, w, &seq, rep);
        return seq.draws;
    }
    uint64_t *ip = (uint64_t *)malloc(sizeof(uint64_t) * (M ? M : 1));
    double *rr = (double *)malloc(sizeof(double) * (M ? M : 1));
    size_t NN = orc_residual_split(M, N, w, ip, rr);
    int kind = scheme == ORC_RESIDUAL ? 0 : (scheme == ORC_RESIDUAL_STRATIFIED ? 1 : 2);
    orc_seq_init(&seq, kind, NN, eng);
    orc_inversion_seq(M, NN, rr, &seq, rep);
    for (size_t i = 0; i < M; ++i) rep[i] += ip[i];
    free(ip);
    free(rr);
    return seq.draws;
}

/* ======================================================================
 * WeightSet -- include/vsmc/core/weight_set.hpp
 *   normalize_log_weight: running max from lw[0], then lw += -max (108-117)
 *   log_weight2weight: w = exp(lw)                                  (369-370)
 *   normalize_weight: coeff = 1/asum(w); w *= coeff; ess = 1/dot(w,w) (119-125)
 * ====================================================================== */
static double orc_post_set_log_weight(size_t N, double *lw, double *w)
{
    double dmax = lw[0];
    for (size_t i = 0; i != N; ++i)
        if (dmax < lw[i]) dmax = lw[i];
    dmax = -dmax;
    for (size_t i = 0; i != N; ++i) lw[i] += dmax;
    for (size_t i = 0; i != N; ++i) w[i] = exp(lw[i]);
    double coeff = 1 / orc_asum(N, w);
    for (size_t i = 0; i != N; ++i) w[i] *= coeff;
    return 1 / orc_dot(N, w, w);
}
double orc_weightset_set_log_weight(size_t N, double *lw, double *w)
{
    return orc_post_set_log_weight(N, lw, w);
}
double orc_weightset_add_log_weight(size_t N, double *lw, const double *inc, double *w)
{
    for (size_t i = 0; i != N; ++i) lw[i] += inc[i];
    return orc_post_set_log_weight(N, lw, w);
}

/* ======================================================================
 * CV particle filter (pf example model, with the constants the workload
 * states).  Transition order follows example/pf/include/pf_smp.hpp:228-233:
 *   x += (sd_pos * z0 + dt * vx); y += (sd_pos * z1 + dt * vy);
 *   vx += sd_vel * z2;            vy += sd_vel * z3.
 * Likelihood: normalised Student-t per coordinate (example/pf/pf_cl.cl:46-58
 * form with log1p, plus the log normaliser llh_const).
 * Normals: one Philox4x32-10 block per particle per step, key = seed, counter
 * {t+1, id_lo, id_hi, 0} (engine ctr set to {t, id, 0}, one increment-then-
 * generate), Box-Muller as in include/vsmc/rng/normal01.h:118-139 with
 * u01_open_closed_32_53: (w0,w1) -> (cos, sin), (w2,w3) -> (cos, sin).
 * ====================================================================== */
void orc_cv_normals(uint64_t seed, uint64_t id, uint64_t t, double z[4])
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)(t + 1), (uint32_t)id, (uint32_t)(id >> 32), 0};
    uint32_t w[4];
    orc_philox4x32_block(ctr, key, w);
    const double two_pi = 6.2831853071795865;
    for (int k = 0; k < 2; ++k) {
        double u1 = orc_u01_open_closed_32_53(w[2 * k]);
        double u2 = orc_u01_open_closed_32_53(w[2 * k + 1]);
        double r = sqrt(-2 * log(u1));
        z[2 * k] = r * cos(two_pi * u2);
        z[2 * k + 1] = r * sin(two_pi * u2);
    }
}

/* One step of the CV model for n particles from the given states (the
 * transition of orc_cv_run; the prior draw at t == 0): st_out and the
 * log-likelihood increments inc. */
void orc_cv_move_many(const orc_cv_param *p, uint64_t seed, size_t t, size_t n,
                      const double *st_in, double ox, double oy, double *st_out, double *inc)
{
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        const size_t i = (size_t)ii;
        double z[4];
        orc_cv_normals(seed, i, t, z);
        double *s = st_out + 4 * i;
        if (t == 0) {
            s[0] = p->prior_sd_pos * z[0];
            s[1] = p->prior_sd_pos * z[1];
            s[2] = p->prior_sd_vel * z[2];
            s[3] = p->prior_sd_vel * z[3];
        } else {
            const double *q = st_in + 4 * i;
            s[0] = q[0] + (p->sd_pos * z[0] + p->dt * q[2]);
            s[1] = q[1] + (p->sd_pos * z[1] + p->dt * q[3]);
            s[2] = q[2] + p->sd_vel * z[2];
            s[3] = q[3] + p->sd_vel * z[3];
        }
        inc[i] = orc_cv_llh(p, s[0], s[1], ox, oy);
    }
}

/* orc_cv_llh of every row (x, y, vx, vy) of st. */
void orc_cv_llh_many(const orc_cv_param *p, size_t n, const double *st, double ox, double oy,
                     double *inc)
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i)
        inc[i] = orc_cv_llh(p, st[4 * i], st[4 * i + 1], ox, oy);
}

double orc_cv_llh(const orc_cv_param *p, double x, double y, double ox, double oy)
{
    double dx = (x - ox) / p->obs_scale;
    double dy = (y - oy) / p->obs_scale;
    double lx = log1p(dx * dx / p->nu);
    double ly = log1p(dy * dy / p->nu);
    return 2 * p->llh_const - 0.5 * (p->nu + 1) * (lx + ly);
}

static double orc_now(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0;
#endif
}

/* The two vSMC SMP backends the CPU path can run the move and the copy on.
 * OpenMP: MoveOMP / StateOMP::copy, one `parallel for` over the particles
 * (include/vsmc/smp/backend_omp.hpp:58-66, 104-117).  std::thread: MoveSTD /
 * StateSTD::copy (include/vsmc/smp/backend_std.hpp:54-62, 98-109) through
 * vsmc::parallel_for / parallel_reduce, which split [0, N) with
 * ThreadNum::partition (include/vsmc/thread/thread_num.hpp:64-89: blocks of
 * ceil(N / threads), or 1 when N < threads) and start one thread per block,
 * joined before the call returns (include/vsmc/thread/parallel_for.hpp:52-63);
 * here POSIX threads do the same. */
typedef struct {
    const orc_cv_param *p;
    double *st, *inc;
    const uint64_t *cf;
    uint64_t seed;
    size_t t;
    double ox, oy;
    int what; /* 0 move, 1 copy */
} orc_cv_work;

static void orc_cv_move_range(const orc_cv_work *w, size_t b, size_t e)
{
    const orc_cv_param *p = w->p;
    for (size_t i = b; i < e; ++i) {
        double z[4];
        orc_cv_normals(w->seed, i, w->t, z);
        double *s = w->st + 4 * i;
        if (w->t == 0) {
            s[0] = p->prior_sd_pos * z[0];
            s[1] = p->prior_sd_pos * z[1];
            s[2] = p->prior_sd_vel * z[2];
            s[3] = p->prior_sd_vel * z[3];
        } else {
            s[0] += p->sd_pos * z[0] + p->dt * s[2];
            s[1] += p->sd_pos * z[1] + p->dt * s[3];
            s[2] += p->sd_vel * z[2];
            s[3] += p->sd_vel * z[3];
        }
        w->inc[i] = orc_cv_llh(p, s[0], s[1], w->ox, w->oy);
    }
}

/* StateMatrix<RowMajor>::copy (core/state_matrix.hpp:285-300): in place is
 * safe in parallel, since cfrp_trans never reads a slot it overwrites */
static void orc_cv_copy_range(const orc_cv_work *w, size_t b, size_t e)
{
    for (size_t to = b; to < e; ++to) {
        const size_t from = (size_t)w->cf[to];
        if (from != to) memcpy(w->st + 4 * to, w->st + 4 * from, sizeof(double) * 4);
    }
}

typedef struct {
    const orc_cv_work *w;
    size_t b, e;
} orc_cv_block;

static void *orc_cv_block_run(void *arg)
{
    const orc_cv_block *k = (const orc_cv_block *)arg;
    if (k->w->what == 0)
        orc_cv_move_range(k->w, k->b, k->e);
    else
        orc_cv_copy_range(k->w, k->b, k->e);
    return NULL;
}

static void orc_cv_parallel(const orc_cv_work *w, size_t N, int threads, int backend)
{
    if (backend == ORC_BACKEND_THREADS && threads > 1 && N > 1) {
        const size_t tn = (size_t)threads;
        const size_t bs = N < tn ? 1 : (N % tn ? N / tn + 1 : N / tn);
        pthread_t tid[ORC_MAX_THREADS];
        orc_cv_block blk[ORC_MAX_THREADS];
        size_t nb = 0;
        for (size_t b = 0; b < N && nb < ORC_MAX_THREADS; b += bs, ++nb) {
            blk[nb].w = w;
            blk[nb].b = b;
            blk[nb].e = b + bs < N ? b + bs : N;
            if (pthread_create(&tid[nb], NULL, orc_cv_block_run, &blk[nb]) != 0) {
                orc_cv_block_run(&blk[nb]);
                tid[nb] = 0;
            }
        }
        for (size_t k = 0; k < nb; ++k)
            if (tid[k]) pthread_join(tid[k], NULL);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ++ii) {
        if (w->what == 0)
            orc_cv_move_range(w, (size_t)ii, (size_t)ii + 1);
        else
            orc_cv_copy_range(w, (size_t)ii, (size_t)ii + 1);
    }
}

static int orc_cv_run_impl(const orc_cv_param *p, size_t N, size_t T, const double *obs,
                           uint64_t seed, uint64_t resample_seed, int scheme, double threshold,
                           double *state_out, double *lw_out, orc_step_record *rec, int threads,
                           int backend, double phase[4])
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    if (threads > ORC_MAX_THREADS) threads = ORC_MAX_THREADS;
    double *st = (double *)malloc(sizeof(double) * 4 * N);
    double *lw = (double *)malloc(sizeof(double) * N);
    double *w = (double *)malloc(sizeof(double) * N);
    double *inc = (double *)malloc(sizeof(double) * N);
    uint64_t *rep = (uint64_t *)malloc(sizeof(uint64_t) * N);
    uint64_t *cf = (uint64_t *)malloc(sizeof(uint64_t) * N);
    if (!st || !lw || !w || !inc || !rep || !cf) {
        free(st); free(lw); free(w); free(inc); free(rep); free(cf);
        return -1;
    }
    orc_tf64_engine reng;
    orc_tf64_seed(&reng, resample_seed);
    double log_z = 0;
    if (phase) phase[0] = phase[1] = phase[2] = phase[3] = 0;

    /* do_reset: set_equal_weight (core/sampler.hpp:758-771) */
    for (size_t i = 0; i < N; ++i) { w[i] = 1.0 / (double)N; lw[i] = 0; }

    for (size_t t = 0; t < T; ++t) {
        double t0 = orc_now();
        const double ox = obs[2 * t], oy = obs[2 * t + 1];
        /* init (t == 0) or move (t >= 1): MoveOMP / MoveSTD over the particles */
        orc_cv_work work = {p, st, inc, cf, seed, t, ox, oy, 0};
        orc_cv_parallel(&work, N, threads, backend);
        double t1 = orc_now();
        /* log-evidence increment: log sum_i W_i exp(inc_i)
         * (example/include/move_smc.hpp:64-71), with a max shift */
        double imax = inc[0];
        for (size_t i = 0; i < N; ++i)
            if (imax < inc[i]) imax = inc[i];
        double sum = 0;
        for (size_t i = 0; i < N; ++i) sum += w[i] * exp(inc[i] - imax);
        log_z += log(sum) + imax;
        double ess;
        if (t == 0) {
            /* do_init: InitializeOMP post_processor -> set_log_weight */
            for (size_t i = 0; i < N; ++i) lw[i] = inc[i];
            ess = orc_post_set_log_weight(N, lw, w);
        } else {
            ess = orc_weightset_add_log_weight(N, lw, inc, w);
        }
        double t2 = orc_now();
        /* do_resample -> Particle::resample (core/particle.hpp:230-250) */
        int resampled = ess < threshold * (double)N;
        size_t draws = 0;
        double t3 = t2;
        if (resampled) {
            draws = orc_resample(scheme, N, N, &reng, w, rep);
            orc_cfrp_trans(N, N, rep, cf);
            t3 = orc_now();
            /* StateOMP::copy / StateSTD::copy: a parallel loop over the slots */
            orc_cv_work cw = {p, st, inc, cf, seed, t, ox, oy, 1};
            orc_cv_parallel(&cw, N, threads, backend);
            /* set_equal_weight (core/weight_set.hpp:96-106) */
            for (size_t i = 0; i < N; ++i) { w[i] = 1.0 / (double)N; lw[i] = 0; }
        }
        double t4 = orc_now();
        /* Monitor at MonitorMCMC: sum_i W_i x_i (core/monitor.hpp:274-302,
         * integrate/is_integrate.hpp:56-80) */
        double mx = 0, my = 0;
        for (size_t i = 0; i < N; ++i) {
            mx = mx + w[i] * st[4 * i];
            my = my + w[i] * st[4 * i + 1];
        }
        if (rec) {
            rec[t].ess = ess;
            rec[t].log_z = log_z;
            rec[t].mean_x = mx;
            rec[t].mean_y = my;
            rec[t].resampled = resampled;
            rec[t].draws = draws;
        }
        if (phase) {
            phase[0] += t1 - t0;
            phase[1] += t2 - t1;
            phase[2] += t3 - t2;
            phase[3] += t4 - t3;
        }
    }
    if (state_out) memcpy(state_out, st, sizeof(double) * 4 * N);
    if (lw_out) memcpy(lw_out, lw, sizeof(double) * N);
    free(st); free(lw); free(w); free(inc); free(rep); free(cf);
    return 0;
}

int orc_cv_run(const orc_cv_param *p, size_t N, size_t T, const double *obs,
               uint64_t seed, uint64_t resample_seed, int scheme, double threshold,
               double *state, double *lw, orc_step_record *rec, int threads)
{
    return orc_cv_run_impl(p, N, T, obs, seed, resample_seed, scheme, threshold, state, lw,
                           rec, threads, ORC_BACKEND_OMP, NULL);
}

int orc_cv_run_timed(const orc_cv_param *p, size_t N, size_t T, const double *obs,
                     uint64_t seed, uint64_t resample_seed, int scheme, double threshold,
                     orc_step_record *rec, int threads, int backend, double phase[4])
{
    return orc_cv_run_impl(p, N, T, obs, seed, resample_seed, scheme, threshold, NULL, NULL,
                           rec, threads, backend, phase);
}

/* ======================================================================
 * The OpenCL-C RNG layer (include/vsmc/rng/urng.h:179-226 with the default
 * cburng4x32 = Threefry4x32-20, urng.h:307-332; include/vsmc/rng/normal01.h:
 * 118-139) and the OpenCL pf example's kernels (example/pf/pf_cl.cl:46-122),
 * the convention user moves follow.  The block function is pinned by the
 * Random123 known-answer vectors; the urng conventions (counter word 0
 * bumped without carry, block words returned last to first) and normal01's
 * draw order are restated from the source: Random123 is absent here, so
 * urng.h and normal01.h cannot be compiled -- parity unpinned for them.
 * ====================================================================== */
void orc_cburng4x32_init(orc_cburng4x32 *r, uint32_t seed)
{
    for (int j = 0; j < 4; ++j) r->key[j] = r->ctr[j] = r->rnd[j] = 0;
    r->key[0] = seed;
    r->remain = 0;
}

uint32_t orc_cburng4x32_rand(orc_cburng4x32 *r)
{
    if (r->remain > 0) return r->rnd[--r->remain];
    ++r->ctr[0];
    orc_threefry4x32_block(r->ctr, r->key, r->rnd);
    r->remain = 3;
    return r->rnd[3];
}

void orc_normal01_4x32_53_init(orc_normal01_53 *n, orc_cburng4x32 *r)
{
    n->u1 = orc_u01_open_closed_32_53(orc_cburng4x32_rand(r));
    n->u2 = orc_u01_open_closed_32_53(orc_cburng4x32_rand(r));
    n->saved = 1;
}

double orc_normal01_4x32_53_rand(orc_normal01_53 *n, orc_cburng4x32 *r)
{
    const double c_2pi = 6.2831853071795865;
    if (n->saved) {
        n->saved = 0;
        return sqrt(-2 * log(n->u1)) * cos(c_2pi * n->u2);
    }
    orc_normal01_4x32_53_init(n, r);
    return sqrt(-2 * log(n->u1)) * sin(c_2pi * n->u2);
}

void orc_cburng4x32_stream(uint32_t seed, size_t n, uint32_t *out)
{
    orc_cburng4x32 r;
    orc_cburng4x32_init(&r, seed);
    for (size_t i = 0; i < n; ++i) out[i] = orc_cburng4x32_rand(&r);
}

void orc_normal01_4x32_53_stream(uint32_t seed, size_t n, double *out)
{
    orc_cburng4x32 r;
    orc_cburng4x32_init(&r, seed);
    orc_normal01_53 z;
    orc_normal01_4x32_53_init(&z, &r);
    for (size_t i = 0; i < n; ++i) out[i] = orc_normal01_4x32_53_rand(&z, &r);
}

/* log_likelihood of pf_cl.cl:46-58 (unnormalised, scale 10, nu 10) */
static double orc_clpf_llh(const double *s, double ox, double oy)
{
    const double scale = 10, nu = 10;
    double lx = scale * (s[0] - ox);
    double ly = scale * (s[1] - oy);
    lx = log1p(lx * lx / nu);
    ly = log1p(ly * ly / nu);
    return -0.5 * (nu + 1) * (lx + ly);
}

/* The OpenCL pf example with fp_type double and NORMAL01_4x32 =
 * normal01_4x32_53: cv_init (pf_cl.cl:60-88) and cv_move (90-122), engine
 * key SEED + i, the counter of slot i kept across iterations (it stays with
 * the slot through a resample), set_log_weight at iteration 0 and
 * add_log_weight after (the MoveCL driver), Particle::resample at threshold
 * (core/particle.hpp:230-250) with the Threefry4x64 resample engine seeded
 * with resample_seed, StateMatrix::copy in index order.  rec: ESS before the
 * resample, log-evidence, resampled, draws (means 0). */
int orc_clpf_run(size_t N, size_t T, const double *obs, uint32_t seed, uint64_t resample_seed,
                 int scheme, double threshold, double *state_out, orc_step_record *rec)
{
    double *st = (double *)malloc(sizeof(double) * 4 * N);
    double *lw = (double *)malloc(sizeof(double) * N);
    double *w = (double *)malloc(sizeof(double) * N);
    double *inc = (double *)malloc(sizeof(double) * N);
    uint32_t *ctr = (uint32_t *)malloc(sizeof(uint32_t) * 4 * N);
    uint64_t *rep = (uint64_t *)malloc(sizeof(uint64_t) * N);
    uint64_t *cf = (uint64_t *)malloc(sizeof(uint64_t) * N);
    if (!st || !lw || !w || !inc || !ctr || !rep || !cf) {
        free(st); free(lw); free(w); free(inc); free(ctr); free(rep); free(cf);
        return -1;
    }
    orc_tf64_engine reng;
    orc_tf64_seed(&reng, resample_seed);
    const double sd_pos = sqrt(0.02), sd_vel = sqrt(0.001), delta = 0.1;
    double log_z = 0;
    for (size_t i = 0; i < N; ++i) w[i] = 1.0 / (double)N;
    for (size_t t = 0; t < T; ++t) {
        const double ox = obs[2 * t], oy = obs[2 * t + 1];
        for (size_t i = 0; i < N; ++i) {
            orc_cburng4x32 r;
            orc_cburng4x32_init(&r, (uint32_t)(seed + i));
            if (t > 0)
                for (int j = 0; j < 4; ++j) r.ctr[j] = ctr[4 * i + j];
            orc_normal01_53 z;
            orc_normal01_4x32_53_init(&z, &r);
            double *s = st + 4 * i;
            if (t == 0) {
                s[0] = orc_normal01_4x32_53_rand(&z, &r) * 2;
                s[1] = orc_normal01_4x32_53_rand(&z, &r) * 2;
                s[2] = orc_normal01_4x32_53_rand(&z, &r) * 1;
                s[3] = orc_normal01_4x32_53_rand(&z, &r) * 1;
            } else {
                s[0] += orc_normal01_4x32_53_rand(&z, &r) * sd_pos + delta * s[2];
                s[1] += orc_normal01_4x32_53_rand(&z, &r) * sd_pos + delta * s[3];
                s[2] += orc_normal01_4x32_53_rand(&z, &r) * sd_vel;
                s[3] += orc_normal01_4x32_53_rand(&z, &r) * sd_vel;
            }
            inc[i] = orc_clpf_llh(s, ox, oy);
            for (int j = 0; j < 4; ++j) ctr[4 * i + j] = r.ctr[j];
        }
        double imax = inc[0];
        for (size_t i = 0; i < N; ++i)
            if (imax < inc[i]) imax = inc[i];
        double sum = 0;
        for (size_t i = 0; i < N; ++i) sum += w[i] * exp(inc[i] - imax);
        log_z += log(sum) + imax;
        double ess;
        if (t == 0) {
            for (size_t i = 0; i < N; ++i) lw[i] = inc[i];
            ess = orc_weightset_set_log_weight(N, lw, w);
        } else {
            ess = orc_weightset_add_log_weight(N, lw, inc, w);
        }
        const int resampled = ess < threshold * (double)N;
        size_t draws = 0;
        if (resampled) {
            draws = orc_resample(scheme, N, N, &reng, w, rep);
            orc_cfrp_trans(N, N, rep, cf);
            for (size_t to = 0; to < N; ++to) {
                const size_t from = cf[to];
                if (from != to) memcpy(st + 4 * to, st + 4 * from, sizeof(double) * 4);
            }
            for (size_t i = 0; i < N; ++i) { w[i] = 1.0 / (double)N; lw[i] = 0; }
        }
        if (rec) {
            rec[t].ess = ess;
            rec[t].log_z = log_z;
            rec[t].mean_x = 0;
            rec[t].mean_y = 0;
            rec[t].resampled = resampled;
            rec[t].draws = draws;
        }
    }
    if (state_out) memcpy(state_out, st, sizeof(double) * 4 * N);
    free(st); free(lw); free(w); free(inc); free(ctr); free(rep); free(cf);
    return 0;
}
