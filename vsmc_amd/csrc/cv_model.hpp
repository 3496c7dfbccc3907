// cv_model.hpp -- the constant-velocity particle model of vSMC's pf example
// (example/pf/include/pf_smp.hpp:96-112, 150-238): state layout, per-particle
// normals and the Student-t log-likelihood, shared by the single-filter and
// the batched kernels.
#pragma once
#include <math.h>

#include "device_util.hpp"
#include "internal.hpp"
#include "rng_core.hpp"

namespace vsmcx {

struct alignas(16) S4f { float x, y, vx, vy; };
struct alignas(32) S4d { double x, y, vx, vy; };

template <typename Real> struct StateOf;
template <> struct StateOf<float> { typedef S4f type; };
template <> struct StateOf<double> { typedef S4d type; };

// four N(0,1) from one Philox4x32-10 block, counter {t+1, id_lo, id_hi, 0}
__device__ __forceinline__ void cv_normals(float z[4], uint32_t k0, uint32_t k1, uint64_t id,
                                           uint64_t t)
{
    u32x4 c;
    c.v[0] = (uint32_t)(t + 1);
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = 0;
    const u32x4 w = philox4x32_10(c, k0, k1);
    box_muller_f32(w.v[0], w.v[1], z[0], z[1]);
    box_muller_f32(w.v[2], w.v[3], z[2], z[3]);
}

__device__ __forceinline__ void cv_normals(double z[4], uint32_t k0, uint32_t k1, uint64_t id,
                                           uint64_t t)
{
    u32x4 c;
    c.v[0] = (uint32_t)(t + 1);
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = 0;
    const u32x4 w = philox4x32_10(c, k0, k1);
    const double two_pi = 6.2831853071795865;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double u1 = u01_oc_53(w.v[2 * j]);
        const double u2 = u01_oc_53(w.v[2 * j + 1]);
        const double r = sqrt(-2 * log(u1));
        z[2 * j] = r * cos(two_pi * u2);
        z[2 * j + 1] = r * sin(two_pi * u2);
    }
}


// One observation (ox, oy) as unevaluated sums hi + lo of two Reals.  In FP32
// a particle's residual x - ox is then (x - hi) - lo, where x - hi is exact for
// x within a factor 2 of hi (Sterbenz): the residual of a particle near the
// observation carries its own rounding only, not the observation's FP32
// rounding, which all particles would share.  In FP64, lo = 0.
template <typename Real>
struct CvObs {
    Real hx, lx, hy, ly;
    __device__ __forceinline__ CvObs(double ox, double oy)
        : hx((Real)ox), lx((Real)(ox - (double)(Real)ox)), hy((Real)oy),
          ly((Real)(oy - (double)(Real)oy))
    {
    }
};

template <typename Real>
struct CvModel {
    Real dt, sd_pos, sd_vel, psd_pos, psd_vel, scale, nu;
    Real inv_scale, inv_nu;
    double c2, half_nu1, half_nu1_ln2;
    __device__ void load(const CvParams &p)
    {
        dt = (Real)p.dt;
        sd_pos = (Real)p.sd_pos;
        sd_vel = (Real)p.sd_vel;
        psd_pos = (Real)p.prior_sd_pos;
        psd_vel = (Real)p.prior_sd_vel;
        scale = (Real)p.obs_scale;
        nu = (Real)p.nu;
        inv_scale = (Real)(1 / p.obs_scale);
        inv_nu = (Real)(1 / p.nu);
        c2 = 2 * p.llh_const;
        half_nu1 = 0.5 * (p.nu + 1);
        half_nu1_ln2 = half_nu1 * 0.69314718055994531;
    }
    // Student-t log-density of the observation, both coordinates
    // (example/pf/include/pf_smp.hpp:96-112, log1p form of example/pf/pf_cl.cl),
    // returned in FP64
    __device__ double llh(Real x, Real y, const CvObs<Real> &o) const;
};

// FP64: the reference's arithmetic (divisions, log1p)
template <>
__device__ __forceinline__ double CvModel<double>::llh(double x, double y,
                                                       const CvObs<double> &o) const
{
    const double dx = (x - o.hx) / scale;
    const double dy = (y - o.hy) / scale;
    const double lx = log1p(dx * dx / nu);
    const double ly = log1p(dy * dy / nu);
    return c2 - half_nu1 * (lx + ly);
}

// FP32: the residuals, squares and log2(1 + s) with s = (1 + a)(1 + b) - 1 in
// FP32 on the hardware log (v_log_f32 is log2); the constant and the factor
// (nu + 1)/2 ln 2 are applied in FP64, so their FP32 roundings, shared by all
// particles, do not enter.  log2(1 + s) = log2(u) s / (u - 1) with u = 1 + s
// rounded (Goldberg, "What every computer scientist should know about
// floating-point arithmetic", 1991, Theorem 4): the correction cancels the
// rounding of u, which for particles near the observation (s ~ 1e-7) is as
// large as s itself; u == 1 leaves log2(1 + s) = s / ln 2.
// The FP32 model's log-likelihood keeps one FP32 term l2 per particle, and the
// FP64 value is c2 - (nu + 1)/2 ln2 l2: a compact log-weight (lw_kind 1)
// stores l2.
__device__ __forceinline__ float cv_llh_l2(const CvModel<float> &md, float x, float y,
                                           const CvObs<float> &o)
{
    const float dx = ((x - o.hx) - o.lx) * md.inv_scale;
    const float dy = ((y - o.hy) - o.ly) * md.inv_scale;
    const float a = dx * dx * md.inv_nu, b = dy * dy * md.inv_nu;
    const float s = __builtin_fmaf(a, b, a + b);
    const float u = 1.0f + s;
    const float d = u - 1.0f;  // exact
    return d == 0.0f ? s * 1.44269504f : __builtin_amdgcn_logf(u) * (s * __builtin_amdgcn_rcpf(d));
}

template <>
__device__ __forceinline__ double CvModel<float>::llh(float x, float y,
                                                      const CvObs<float> &o) const
{
    return c2 - half_nu1_ln2 * (double)cv_llh_l2(*this, x, y, o);
}

// x += sd_pos z0 + dt vx (old velocity), v += sd_vel z; the prior at t = 0
// (example/pf/include/pf_smp.hpp:150-164, 207-238)
template <typename Real, typename S4>
__device__ __forceinline__ void cv_advance(const CvModel<Real> &md, S4 &s, const Real z[4],
                                           uint64_t t)
{
    if (t == 0) {
        s.x = md.psd_pos * z[0];
        s.y = md.psd_pos * z[1];
        s.vx = md.psd_vel * z[2];
        s.vy = md.psd_vel * z[3];
    } else {
        s.x += md.sd_pos * z[0] + md.dt * s.vx;
        s.y += md.sd_pos * z[1] + md.dt * s.vy;
        s.vx += md.sd_vel * z[2];
        s.vy += md.sd_vel * z[3];
    }
}

// Per-workgroup sums of a step: e^(lw-M), e^(2(lw-M)), and the monitor sums.
struct Part5 { double se, se2, mx, my; };
__device__ __forceinline__ Part5 lane_xor(const Part5 &v, int m)
{
    return Part5{lane_xor(v.se, m), lane_xor(v.se2, m), lane_xor(v.mx, m), lane_xor(v.my, m)};
}
struct Part5Add {
    __device__ Part5 operator()(const Part5 &a, const Part5 &b) const
    {
        return Part5{a.se + b.se, a.se2 + b.se2, a.mx + b.mx, a.my + b.my};
    }
};

// Online log-sum-exp of FP64 log-weights: running max m, sum e^(lw-m) and
// sum e^(2(lw-m)), one exp per value.
__device__ __forceinline__ void lse_add(double &m, double &se, double &se2, double lw)
{
    if (lw > m) {
        const double r = exp(m - lw);
        se = se * r + 1;
        se2 = se2 * r * r + 1;
        m = lw;
    } else {
        const double e = exp(lw - m);
        se += e;
        se2 += e * e;
    }
}

// After a step's reductions (one thread): log-sum-exp, ESS, log-evidence
// increment LSE_new - LSE_old (LSE_old = log N after a resample or at the
// start), the resampling decision (ess < threshold * N, core/particle.hpp:232)
// and the history row; the previous step's MonitorMCMC mean goes to row t-1.
__device__ __forceinline__ void cv_step_scalars(FilterCtl *ctl, double M, double se, double se2,
                                                double mx, double my, StepRow *hist, uint64_t t)
{
    const double n = (double)ctl->n;
    const double lse_old = (t == 0 || ctl->prev_resampled) ? log(n) : ctl->lse;
    const double lse = M + log(se);
    const double logz = (t == 0 ? 0.0 : ctl->logz) + (lse - lse_old);
    const double ess = se * se / se2;
    int res = ess < ctl->threshold * n;
    if (!(se > 0) || !isfinite(lse) || !isfinite(ess)) {
        atomicOr(&ctl->err, (int)kErrDegenerate);
        res = 0;
    }
    ctl->m = M;
    ctl->coeff = 1 / se;
    ctl->lse = lse;
    ctl->logz = logz;
    ctl->ess = ess;
    ctl->resample = res;
    ctl->t = t;
    ctl->prev_resampled = 0;
    ctl->lw_kind = ctl->lw_store_kind;
    ctl->r_int = 0;
    ctl->draws = 0;
    if (hist) {
        StepRow r;
        r.ess = ess;
        r.logz = logz;
        r.mean_x = 0;
        r.mean_y = 0;
        r.resampled = res;
        r.pad = 0;
        r.draws = 0;
        hist[t] = r;
        if (t > 0) {
            hist[t - 1].mean_x = mx;
            hist[t - 1].mean_y = my;
        }
    }
}

}  // namespace vsmcx
