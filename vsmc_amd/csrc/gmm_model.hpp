// gmm_model.hpp -- the Gaussian-mixture model of the tempered SMC sampler:
// prior, likelihood, initial draws and the three random-walk MH moves of one
// particle, in the FP64 arithmetic of the reference (fused FP64 path) and in
// FP32 with hardware transcendentals (mixed path).
//
// Model (example/gmm/include/gmm_state.hpp:73-116): K components with means
// mu_d, precisions lambda_d and weights w_d;
//   log prior  = sum_d -(mu_d - mu0)^2 / (2 sd0^2) + (shape0 - 1) log lambda_d - lambda_d / scale0
//   log lik    = sum_i log(sum_d w_d sqrt(lambda_d) exp(-lambda_d (x_i - mu_d)^2 / 2) + 1e-13)
//   log target = log prior + alpha * log lik.
// Moves (example/gmm/include/gmm_move.hpp:35-121): mu_d += sd_mu z_d;
// lambda_d *= exp(sd_lambda z_d) with Jacobian sum_d log(lambda'_d / lambda_d);
// logits log(w_d / w_{K-1}) += sd_w z_d (d < K-1) with the logit Jacobian of
// example/gmm/include/gmm_param.hpp:221-248; accept iff p >= log u
// (gmm_param.hpp:250-324).
//
// Random numbers: Philox4x32-10 keyed by the seed, counter
// {iteration, id_lo, id_hi, block}.  Initial draws (iteration 0) read the
// word stream of blocks 0..6: words 0..7 Box-Muller pairs for mu, words 8..19
// two exponentials per lambda (Gamma(2, scale0)), words 20..25 one exponential
// per weight (Dirichlet(1) by normalised Gamma(1, 1)).  MH move b of sweep s
// at iteration t reads blocks 8 + 2 (3 s + b) and the next: words 0..5 three
// Box-Muller pairs, word 7 the uniform of log u.
#pragma once
#include "device_util.hpp"
#include "launch.hpp"
#include "rng_core.hpp"

namespace vsmcx {

__device__ __forceinline__ u32x4 gmm_block(const GmmModel &g, uint64_t id, uint32_t iter,
                                           uint32_t blk)
{
    u32x4 c;
    c.v[0] = iter;
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = blk;
    return philox4x32_10(c, (uint32_t)g.seed, (uint32_t)(g.seed >> 32));
}

// ---- Box-Muller and exponentials in each precision --------------------------------
__device__ __forceinline__ void gmm_bm(uint32_t a, uint32_t b, double &z0, double &z1)
{
    const double two_pi = 6.2831853071795865;
    const double u1 = u01_oc_53(a), u2 = u01_oc_53(b);
    const double r = sqrt(-2 * log(u1));
    z0 = r * cos(two_pi * u2);
    z1 = r * sin(two_pi * u2);
}
__device__ __forceinline__ void gmm_bm(uint32_t a, uint32_t b, float &z0, float &z1)
{
    box_muller_f32(a, b, z0, z1);
}
__device__ __forceinline__ double gmm_log_u01(uint32_t a, double) { return log(u01_oc_53(a)); }
__device__ __forceinline__ double gmm_log_u01(uint32_t a, float)
{
    return (double)(0.69314718f * __builtin_amdgcn_logf(u01_oo_24(a)));
}
__device__ __forceinline__ double gmm_log(double x) { return log(x); }
__device__ __forceinline__ float gmm_log(float x) { return 0.69314718f * __builtin_amdgcn_logf(x); }
__device__ __forceinline__ double gmm_exp(double x) { return exp(x); }
__device__ __forceinline__ float gmm_exp(float x) { return __builtin_amdgcn_exp2f(1.44269504f * x); }

// ---- prior ---------------------------------------------------------------------------
template <int K, typename Real>
__device__ __forceinline__ Real gmm_log_prior(const GmmModel &g, const Real *mu, const Real *lam)
{
    const Real mu0 = (Real)g.mu0, sd0 = (Real)g.sd0, sh1 = (Real)(g.shape0 - 1),
               scale0 = (Real)g.scale0;
    Real lp = 0;
#pragma unroll
    for (int d = 0; d < K; ++d) {
        const Real resid = mu[d] - mu0;
        lp += (Real)-0.5 * (resid * resid) / (sd0 * sd0);
        lp += sh1 * gmm_log(lam[d]) - lam[d] / scale0;
    }
    return lp;
}

// ---- likelihood --------------------------------------------------------------------
// FP64: the reference's expression, point by point.
template <int K>
__device__ __forceinline__ double gmm_log_lik(const GmmModel &g, const double *mu,
                                              const double *lam, const double *w)
{
    double hl[K];
#pragma unroll
    for (int d = 0; d < K; ++d) hl[d] = 0.5 * log(lam[d]);
    const double *__restrict__ x = g.xd;
    double ll = 0;
#pragma unroll 1
    for (uint32_t i = 0; i < g.n; ++i) {
        const double xi = x[i];
        double lli = 0;
#pragma unroll
        for (int d = 0; d < K; ++d) {
            const double r = xi - mu[d];
            lli += w[d] * exp(hl[d] - 0.5 * lam[d] * r * r);
        }
        ll += log(lli + 1e-13);
    }
    return ll;
}

// FP32: each component term is 2^(C_d - t^2) with t = (x - mu_d) s_d,
// s_d = sqrt(lambda_d log2(e) / 2), C_d = log2(w_d sqrt(lambda_d)): t is one
// FMA x s_d + (-mu_d s_d) with the wave-uniform x as a scalar operand, the
// exponent one more FMA, then one v_exp_f32 and the accumulation -- four VALU
// instructions per component and point, each rounding relative to the
// quantity it produces.  The mixtures of two consecutive points are
// multiplied before one v_log_f32 (each factor lies in
// [1e-13, sum_d w_d sqrt(lambda_d)], so the product stays a normal float),
// and the log2 values are summed in FP32 over 32 points at a time, then in
// FP64.
template <int K>
__device__ __forceinline__ double gmm_log_lik(const GmmModel &g, const float *mu, const float *lam,
                                              const float *w)
{
    float S[K], M[K], Cc[K];
#pragma unroll
    for (int d = 0; d < K; ++d) {
        S[d] = __builtin_amdgcn_sqrtf(0.72134752f * lam[d]);
        M[d] = -mu[d] * S[d];
        Cc[d] = __builtin_amdgcn_logf(w[d]) + 0.5f * __builtin_amdgcn_logf(lam[d]);
    }
    const float *__restrict__ x = g.xf;
    const uint32_t n = g.n;
    double acc = 0;
    float part = 0;
    uint32_t i = 0;
#pragma unroll 1
    for (; i + 32 <= n; i += 32) {
#pragma unroll 4
        for (int j = 0; j < 32; j += 2) {
            const float x0 = x[i + j], x1 = x[i + j + 1];
            float m0 = 1e-13f, m1 = 1e-13f;
#pragma unroll
            for (int d = 0; d < K; ++d) {
                const float t0 = __builtin_fmaf(x0, S[d], M[d]);
                const float t1 = __builtin_fmaf(x1, S[d], M[d]);
                m0 += __builtin_amdgcn_exp2f(__builtin_fmaf(-t0, t0, Cc[d]));
                m1 += __builtin_amdgcn_exp2f(__builtin_fmaf(-t1, t1, Cc[d]));
            }
            part += __builtin_amdgcn_logf(m0 * m1);
        }
        acc += (double)part;
        part = 0;
    }
#pragma unroll 1
    for (; i < n; ++i) {
        const float x0 = x[i];
        float m0 = 1e-13f;
#pragma unroll
        for (int d = 0; d < K; ++d) {
            const float t0 = __builtin_fmaf(x0, S[d], M[d]);
            m0 += __builtin_amdgcn_exp2f(__builtin_fmaf(-t0, t0, Cc[d]));
        }
        part += __builtin_amdgcn_logf(m0);
    }
    acc += (double)part;
    return acc * 0.69314718055994531;
}

// ---- initial draws (gmm_init.hpp:40-70) ---------------------------------------------
template <int K, typename Real>
__device__ __forceinline__ void gmm_init_draw(const GmmModel &g, uint64_t id, Real *mu, Real *lam,
                                              Real *w)
{
    uint32_t wd[28];
#pragma unroll
    for (int b = 0; b < 7; ++b) {
        const u32x4 r = gmm_block(g, id, 0, (uint32_t)b);
#pragma unroll
        for (int j = 0; j < 4; ++j) wd[4 * b + j] = r.v[j];
    }
    Real z[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) gmm_bm(wd[2 * k], wd[2 * k + 1], z[2 * k], z[2 * k + 1]);
    Real sum = 0;
    const Real one = 1;
#pragma unroll
    for (int d = 0; d < K; ++d) {
        mu[d] = (Real)g.mu0 + (Real)g.sd0 * z[d];
        lam[d] = (Real)g.scale0 * (-(Real)gmm_log_u01(wd[8 + 2 * d], one) -
                                   (Real)gmm_log_u01(wd[9 + 2 * d], one));
        w[d] = -(Real)gmm_log_u01(wd[20 + d], one);
        sum += w[d];
    }
#pragma unroll
    for (int d = 0; d < K; ++d) w[d] /= sum;
}

// logit part of the weight Jacobian (gmm_param.hpp:221-248)
template <int K, typename Real>
__device__ __forceinline__ Real gmm_logit_part(const Real *w)
{
    Real sum_lw = 1, sum_llw = 0;
#pragma unroll
    for (int d = 0; d < K - 1; ++d) {
        const Real r = w[d] / w[K - 1];
        sum_lw += r;
        sum_llw += gmm_log(r);
    }
    return sum_llw - (Real)K * gmm_log(sum_lw);
}

// One particle's state during the MH sweeps.
template <int K, typename Real>
struct GmmParticle {
    Real mu[K], lam[K], w[K];
    double llh;  // log likelihood
    Real lp;     // log prior
};

// One MH move b (0 mu, 1 lambda, 2 weight) of sweep s; returns 1 if accepted.
template <int K, typename Real>
__device__ __forceinline__ int gmm_mh_move(const GmmModel &g, GmmParticle<K, Real> &p, int b,
                                           uint64_t id, uint32_t iter, uint32_t s, double alpha,
                                           const double *sd)
{
    const uint32_t blk = 8 + 2 * (3 * s + (uint32_t)b);
    const u32x4 r0 = gmm_block(g, id, iter, blk), r1 = gmm_block(g, id, iter, blk + 1);
    const uint32_t wd[8] = {r0.v[0], r0.v[1], r0.v[2], r0.v[3], r1.v[0], r1.v[1], r1.v[2], r1.v[3]};
    const Real one = 1;
    Real z[6];
    const int need = b == 2 ? K - 1 : K;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (2 * k < need) gmm_bm(wd[2 * k], wd[2 * k + 1], z[2 * k], z[2 * k + 1]);
    const double log_u = gmm_log_u01(wd[7], one);
    Real mu[K], lam[K], w[K];
#pragma unroll
    for (int d = 0; d < K; ++d) {
        mu[d] = p.mu[d];
        lam[d] = p.lam[d];
        w[d] = p.w[d];
    }
    Real jac = 0;
    if (b == 0) {
        const Real sdm = (Real)sd[0];
#pragma unroll
        for (int d = 0; d < K; ++d) mu[d] += sdm * z[d];
    } else if (b == 1) {
        const Real sdl = (Real)sd[1];
#pragma unroll
        for (int d = 0; d < K; ++d) lam[d] *= gmm_exp(sdl * z[d]);
#pragma unroll
        for (int d = 0; d < K; ++d) jac += gmm_log(lam[d]) - gmm_log(p.lam[d]);
    } else {
        const Real sdw = (Real)sd[2];
        Real sum_weight = 1;
#pragma unroll
        for (int d = 0; d < K - 1; ++d) {
            w[d] = gmm_log(w[d] / w[K - 1]);
            w[d] += sdw * z[d];
            w[d] = gmm_exp(w[d]);
            sum_weight += w[d];
        }
        w[K - 1] = 1;
#pragma unroll
        for (int d = 0; d < K; ++d) w[d] /= sum_weight;
        jac = gmm_logit_part<K>(w) - gmm_logit_part<K>(p.w);
    }
    const Real lp = gmm_log_prior<K>(g, mu, lam);
    const double ll = gmm_log_lik<K>(g, mu, lam, w);
    const double pr = ((double)lp + alpha * ll) - ((double)p.lp + alpha * p.llh) + (double)jac;
    if (pr < log_u) return 0;
#pragma unroll
    for (int d = 0; d < K; ++d) {
        p.mu[d] = mu[d];
        p.lam[d] = lam[d];
        p.w[d] = w[d];
    }
    p.llh = ll;
    p.lp = lp;
    return 1;
}

}  // namespace vsmcx
