// kernels_cv.hip -- constant-velocity particle filter (example/pf model).
//
// k_cv_move fuses, per particle, what vSMC spreads over MoveOMP's parallel
// loop and WeightSet's sequential passes:
//   * ancestor gather of the previous resample (StateMatrix::copy,
//     include/vsmc/core/state_matrix.hpp:285-300) folded into the state read,
//   * one Philox4x32-10 block per particle per step, Box-Muller normals
//     (the design of include/vsmc/rng/normal01.h:118-139),
//   * the transition and Student-t log-likelihood of
//     example/pf/include/pf_smp.hpp:96-112,207-238 (old velocity in the
//     position update, log1p form of example/pf/pf_cl.cl:46-58),
//   * lw += inc in FP64 (WeightSet::add_log_weight, core/weight_set.hpp:324-331),
//   * an online log-sum-exp over the new log-weights (max, sum e, sum e^2)
//     with one FP64 exp per particle, replacing normalize_log_weight /
//     vExp / asum / dot (core/weight_set.hpp:108-125),
//   * the previous step's MonitorMCMC weighted mean (core/monitor.hpp:274-302)
//     accumulated from the state it reads anyway.
// State is AoS: 4 x FP32 (16 B) in the mixed pipeline, 4 x FP64 (32 B) in
// the FP64 pipelines; each wave load/store moves one contiguous 1-2 KiB span.
#include <math.h>

#include "device_util.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "cv_model.hpp"
#include "rng_core.hpp"

namespace vsmcx {

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_cv_move(
    FilterCtl *ctl, const typename StateOf<Real>::type *__restrict__ st_in,
    typename StateOf<Real>::type *__restrict__ st_out, const uint32_t *__restrict__ anc,
    double *__restrict__ lw, const double *__restrict__ obs, CvParams p, uint32_t k0,
    uint32_t k1, uint64_t t, uint64_t N, double *__restrict__ partials)
{
    typedef typename StateOf<Real>::type S4;
    __shared__ double s_red[kBlock / kWave];
    __shared__ Part5 s_p5[kBlock / kWave];
    CvModel<Real> md;
    md.load(p);
    const int prev = t > 0 ? ctl->prev_resampled : 0;
    const double m_old = ctl->m, c_old = ctl->coeff;
    const double invN = 1.0 / (double)N;
    const Real ox = (Real)obs[2 * t], oy = (Real)obs[2 * t + 1];
    const uint64_t base = blockIdx.x * (uint64_t)kMoveTile;

    // Loads first, for all of the thread's particles (ancestor indices, then
    // the gathered states and old log-weights), so their latencies overlap;
    // then the arithmetic and the stores.
    uint64_t src[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        uint64_t j = (t > 0 && prev && i < N) ? (uint64_t)anc[i] : i;
        src[k] = j < N ? j : (i < N ? i : 0);  // never read outside the state buffer
    }
    S4 s[kMoveItems];
    double lwo[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        lwo[k] = 0;
        if (t > 0 && i < N) {
            s[k] = st_in[src[k]];
            if (!prev) lwo[k] = lw[i];
        }
    }

    double m = -INFINITY, se = 0, se2 = 0, mx = 0, my = 0;
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i >= N) continue;
        S4 &sk = s[k];
        if (t > 0) {
            const double w = prev ? invN : exp(lwo[k] - m_old) * c_old;
            mx += w * (double)sk.x;
            my += w * (double)sk.y;
        }
        Real z[4];
        cv_normals(z, k0, k1, i, t);
        cv_advance(md, sk, z, t);
        const Real inc = md.llh(sk.x, sk.y, ox, oy);
        const double lwn = lwo[k] + (double)inc;
        st_out[i] = sk;
        lw[i] = lwn;
        lse_add(m, se, se2, lwn);
    }
    const double M = block_reduce<kBlock>(m, MaxOp(), s_red);
    const double sc = (m == -INFINITY) ? 0.0 : exp(m - M);
    Part5 v{se * sc, se2 * sc * sc, mx, my};
    v = block_reduce<kBlock>(v, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        double *o = partials + 5 * (uint64_t)blockIdx.x;
        o[0] = M;
        o[1] = v.se;
        o[2] = v.se2;
        o[3] = v.mx;
        o[4] = v.my;
    }
}

// One workgroup: combine tile partials, ESS, log-evidence, resample decision.
__global__ __launch_bounds__(1024) void k_cv_reduce(FilterCtl *ctl, const double *partials,
                                                    uint64_t ntiles, StepRow *hist, uint64_t t)
{
    __shared__ double s_red[1024 / kWave];
    __shared__ Part5 s_p5[1024 / kWave];
    double m = -INFINITY;
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) {
        const double v = partials[5 * j];
        m = m < v ? v : m;
    }
    const double M = block_reduce<1024>(m, MaxOp(), s_red);
    Part5 acc{0, 0, 0, 0};
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) {
        const double *q = partials + 5 * j;
        const double sc = q[0] == -INFINITY ? 0.0 : exp(q[0] - M);
        acc.se += q[1] * sc;
        acc.se2 += q[2] * sc * sc;
        acc.mx += q[3];
        acc.my += q[4];
    }
    acc = block_reduce<1024>(acc, Part5Add(), s_p5);
    if (threadIdx.x == 0) cv_step_scalars(ctl, M, acc.se, acc.se2, acc.mx, acc.my, hist, t);
}

// Monitor of the last step: sum_i W_i (x_i, y_i) over the post-resample state.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_cv_monitor(FilterCtl *ctl,
                                                       const typename StateOf<Real>::type *st,
                                                       const uint32_t *anc, const double *lw,
                                                       uint64_t N, double *partials)
{
    __shared__ Part5 s_p5[kBlock / kWave];
    const int prev = ctl->prev_resampled;
    const double m = ctl->m, c = ctl->coeff, invN = 1.0 / (double)N;
    const uint64_t base = blockIdx.x * (uint64_t)kMoveTile;
    Part5 v{0, 0, 0, 0};
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i >= N) continue;
        const auto s = st[prev ? (uint64_t)anc[i] : i];
        const double w = prev ? invN : exp(lw[i] - m) * c;
        v.mx += w * (double)s.x;
        v.my += w * (double)s.y;
    }
    v = block_reduce<kBlock>(v, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        partials[5 * (uint64_t)blockIdx.x + 3] = v.mx;
        partials[5 * (uint64_t)blockIdx.x + 4] = v.my;
    }
}

__global__ __launch_bounds__(1024) void k_cv_monitor_sum(FilterCtl *ctl, const double *partials,
                                                         uint64_t ntiles, StepRow *hist)
{
    __shared__ Part5 s_p5[1024 / kWave];
    Part5 acc{0, 0, 0, 0};
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) {
        acc.mx += partials[5 * j + 3];
        acc.my += partials[5 * j + 4];
    }
    acc = block_reduce<1024>(acc, Part5Add(), s_p5);
    if (threadIdx.x == 0 && hist) {
        hist[ctl->t].mean_x = acc.mx;
        hist[ctl->t].mean_y = acc.my;
    }
}

// W_i exactly as the resampler computes it (XWeight in kernels_fold.hip)
__global__ void k_cv_weights(const FilterCtl *ctl, const double *lw, uint64_t N, double *w)
{
    const double m = ctl->m, c = ctl->coeff;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = exp(lw[i] - m) * c;
}

template <typename S4>
__global__ void k_cv_gather(const FilterCtl *ctl, const S4 *in, S4 *out, const uint32_t *anc,
                            uint64_t N)
{
    const int prev = ctl->prev_resampled;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[prev ? (uint64_t)anc[i] : i];
}

static unsigned tiles_of(uint64_t N) { return (unsigned)((N + kMoveTile - 1) / kMoveTile); }

void launch_cv_move(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                    const uint32_t *anc, double *lw, const double *obs, const CvParams &p,
                    uint64_t seed, uint64_t t, uint64_t N, double *partials, hipStream_t st)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (fp64)
        hipLaunchKernelGGL(k_cv_move<double>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4d *)st_in, (S4d *)st_out, anc, lw, obs, p, k0, k1, t, N,
                           partials);
    else
        hipLaunchKernelGGL(k_cv_move<float>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4f *)st_in, (S4f *)st_out, anc, lw, obs, p, k0, k1, t, N,
                           partials);
}

void launch_cv_reduce(FilterCtl *ctl, const double *partials, uint64_t ntiles, StepRow *hist,
                      uint64_t t, hipStream_t st)
{
    hipLaunchKernelGGL(k_cv_reduce, dim3(1), dim3(1024), 0, st, ctl, partials, ntiles, hist, t);
}

void launch_cv_monitor(bool fp64, FilterCtl *ctl, const void *stp, const uint32_t *anc,
                       const double *lw, uint64_t N, double *partials, StepRow *hist,
                       hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_cv_monitor<double>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4d *)stp, anc, lw, N, partials);
    else
        hipLaunchKernelGGL(k_cv_monitor<float>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4f *)stp, anc, lw, N, partials);
    hipLaunchKernelGGL(k_cv_monitor_sum, dim3(1), dim3(1024), 0, st, ctl, partials,
                       (uint64_t)tiles_of(N), hist);
}

void launch_cv_weights(FilterCtl *ctl, const double *lw, uint64_t N, double *w, hipStream_t st)
{
    hipLaunchKernelGGL(k_cv_weights, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl, lw, N, w);
}

void launch_cv_gather(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                      const uint32_t *anc, uint64_t N, hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_cv_gather<S4d>, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl,
                           (const S4d *)st_in, (S4d *)st_out, anc, N);
    else
        hipLaunchKernelGGL(k_cv_gather<S4f>, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl,
                           (const S4f *)st_in, (S4f *)st_out, anc, N);
}

}  // namespace vsmcx
