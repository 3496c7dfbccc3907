// kernels_plain.hip -- the un-fused FP64 pipeline, one kernel per vSMC operation.
//
// This is the baseline the fused pipelines are measured against: it keeps
// vSMC's per-step operation sequence and its in-place state, each operation a
// separate pass over HBM:
//   move + llh            MoveOMP::operator() (include/vsmc/smp/backend_omp.hpp:104-117)
//   log-evidence          example/include/move_smc.hpp:64-71
//   add_log_weight        core/weight_set.hpp:324-331
//   normalize_log_weight  max, subtract (core/weight_set.hpp:108-117)
//   log_weight2weight     vExp (core/weight_set.hpp:369-370)
//   normalize_weight      asum, scal, dot (core/weight_set.hpp:119-125)
//   resample              Particle::resample (core/particle.hpp:230-250)
//   copy                  StateMatrix::copy in place (core/state_matrix.hpp:285-300)
//   set_equal_weight      core/weight_set.hpp:96-106
//   monitor               Monitor::eval at MonitorMCMC (core/monitor.hpp:274-302)
// Reductions are two-pass (tile partials, then one workgroup) so results do
// not depend on scheduling.
#include <math.h>

#include "device_util.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "rng_core.hpp"

namespace vsmcx {

struct alignas(32) P4d { double x, y, vx, vy; };

enum PlainSlot : int { kMaxInc = 0, kSumW = 1, kMaxLw = 2, kAsum = 3, kDot = 4, kMx = 5, kMy = 6 };

__global__ __launch_bounds__(kBlock) void k_plain_move(P4d *st, double *inc, const double *obs,
                                                       CvParams p, uint32_t k0, uint32_t k1,
                                                       uint64_t t, uint64_t N)
{
    const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
    if (i >= N) return;
    u32x4 c;
    c.v[0] = (uint32_t)(t + 1);
    c.v[1] = (uint32_t)i;
    c.v[2] = (uint32_t)(i >> 32);
    c.v[3] = 0;
    const u32x4 w = philox4x32_10(c, k0, k1);
    const double two_pi = 6.2831853071795865;
    double z[4];
    for (int j = 0; j < 2; ++j) {
        const double u1 = u01_oc_53(w.v[2 * j]), u2 = u01_oc_53(w.v[2 * j + 1]);
        const double r = sqrt(-2 * log(u1));
        z[2 * j] = r * cos(two_pi * u2);
        z[2 * j + 1] = r * sin(two_pi * u2);
    }
    P4d s;
    if (t == 0) {
        s.x = p.prior_sd_pos * z[0];
        s.y = p.prior_sd_pos * z[1];
        s.vx = p.prior_sd_vel * z[2];
        s.vy = p.prior_sd_vel * z[3];
    } else {
        s = st[i];
        s.x += p.sd_pos * z[0] + p.dt * s.vx;
        s.y += p.sd_pos * z[1] + p.dt * s.vy;
        s.vx += p.sd_vel * z[2];
        s.vy += p.sd_vel * z[3];
    }
    st[i] = s;
    const double dx = (s.x - obs[2 * t]) / p.obs_scale, dy = (s.y - obs[2 * t + 1]) / p.obs_scale;
    inc[i] = 2 * p.llh_const - 0.5 * (p.nu + 1) * (log1p(dx * dx / p.nu) + log1p(dy * dy / p.nu));
}

// --- two-pass reductions -----------------------------------------------------
// op 0: max(a); 1: sum(a); 2: sum(a*b); 3: sum(w * exp(a - s[kMaxInc]))
// 4: sum(w * state.x) and sum(w * state.y)
template <int OP>
__global__ __launch_bounds__(kBlock) void k_plain_red(const double *a, const double *b,
                                                      const P4d *st, const double *scal,
                                                      uint64_t N, double *part)
{
    __shared__ double s_red[kBlock / kWave];
    double acc = OP == 0 ? -INFINITY : 0.0, acc2 = 0.0;
    const double c = OP == 3 ? scal[kMaxInc] : 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * kBlock) {
        if (OP == 0) acc = acc < a[i] ? a[i] : acc;
        if (OP == 1) acc += a[i];
        if (OP == 2) acc += a[i] * b[i];
        if (OP == 3) acc += b[i] * exp(a[i] - c);
        if (OP == 4) {
            acc += b[i] * st[i].x;
            acc2 += b[i] * st[i].y;
        }
    }
    if (OP == 0)
        acc = block_reduce<kBlock>(acc, MaxOp(), s_red);
    else
        acc = block_reduce<kBlock>(acc, AddOp(), s_red);
    if (OP == 4) acc2 = block_reduce<kBlock>(acc2, AddOp(), s_red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = acc;
        part[2 * blockIdx.x + 1] = acc2;
    }
}

template <int OP>
__global__ __launch_bounds__(1024) void k_plain_red_final(const double *part, uint64_t nparts,
                                                          double *scal, int slot)
{
    __shared__ double s_red[1024 / kWave];
    double acc = OP == 0 ? -INFINITY : 0.0, acc2 = 0.0;
    for (uint64_t j = threadIdx.x; j < nparts; j += 1024) {
        if (OP == 0)
            acc = acc < part[2 * j] ? part[2 * j] : acc;
        else
            acc += part[2 * j];
        acc2 += part[2 * j + 1];
    }
    if (OP == 0)
        acc = block_reduce<1024>(acc, MaxOp(), s_red);
    else
        acc = block_reduce<1024>(acc, AddOp(), s_red);
    acc2 = block_reduce<1024>(acc2, AddOp(), s_red);
    if (threadIdx.x == 0) {
        scal[slot] = acc;
        if (OP == 4) scal[slot + 1] = acc2;
    }
}

// --- element-wise passes -------------------------------------------------------
__global__ void k_plain_add(double *lw, const double *inc, uint64_t N, int set)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) lw[i] = set ? inc[i] : lw[i] + inc[i];
}
__global__ void k_plain_sub(double *lw, const double *scal, uint64_t N)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) lw[i] += -scal[kMaxLw];
}
__global__ void k_plain_exp(const double *lw, double *w, uint64_t N)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) w[i] = exp(lw[i]);
}
__global__ void k_plain_scal(double *w, const double *scal, uint64_t N)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) w[i] *= 1 / scal[kAsum];
}

// ESS, log-evidence, decision; the FilterCtl fields the resampler reads
__global__ void k_plain_decide(FilterCtl *ctl, const double *scal, StepRow *hist, uint64_t t)
{
    if (threadIdx.x != 0) return;
    const double ess = 1 / scal[kDot];
    const double logz = (t == 0 ? 0.0 : ctl->logz) + log(scal[kSumW]) + scal[kMaxInc];
    int res = ess < ctl->threshold * (double)ctl->n;
    if (!(scal[kAsum] > 0) || !isfinite(ess)) {
        atomicOr(&ctl->err, (int)kErrDegenerate);
        res = 0;
    }
    ctl->ess = ess;
    ctl->logz = logz;
    ctl->resample = res;
    ctl->t = t;
    ctl->prev_resampled = 0;
    ctl->r_int = 0;
    ctl->draws = 0;
    if (hist) {
        StepRow r{};
        r.ess = ess;
        r.logz = logz;
        r.resampled = res;
        hist[t] = r;
    }
}

__global__ void k_plain_copy(const FilterCtl *ctl, P4d *st, const uint32_t *anc, uint64_t N)
{
    if (!ctl->resample) return;
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N && anc[i] != i) st[i] = st[anc[i]];
}

__global__ void k_plain_reset(const FilterCtl *ctl, double *w, double *lw, uint64_t N)
{
    if (!ctl->resample) return;
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) {
        w[i] = 1 / (double)N;
        lw[i] = 0;
    }
}

__global__ void k_plain_init_w(double *w, uint64_t N)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) w[i] = 1 / (double)N;
}

__global__ void k_plain_mean(const double *scal, StepRow *hist, uint64_t t)
{
    if (threadIdx.x == 0 && hist) {
        hist[t].mean_x = scal[kMx];
        hist[t].mean_y = scal[kMy];
    }
}

static unsigned g256(uint64_t N) { return (unsigned)((N + 255) / 256); }
static unsigned gred(uint64_t N)
{
    const uint64_t g = (N + kBlock * 16 - 1) / (kBlock * 16);
    return (unsigned)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

template <int OP>
static void reduce(const double *a, const double *b, const P4d *st, double *scal, int slot,
                   uint64_t N, double *part, hipStream_t s)
{
    const unsigned g = gred(N);
    hipLaunchKernelGGL(k_plain_red<OP>, dim3(g), dim3(kBlock), 0, s, a, b, st, scal, N, part);
    hipLaunchKernelGGL(k_plain_red_final<OP>, dim3(1), dim3(1024), 0, s, part, (uint64_t)g, scal,
                       slot);
}

void launch_plain_init_weights(double *w, uint64_t N, hipStream_t s)
{
    hipLaunchKernelGGL(k_plain_init_w, dim3(g256(N)), dim3(256), 0, s, w, N);
}

void launch_plain_move(void *st, double *inc, const double *obs, const CvParams &p, uint64_t seed,
                       uint64_t t, uint64_t N, hipStream_t s)
{
    hipLaunchKernelGGL(k_plain_move, dim3(g256(N)), dim3(kBlock), 0, s, (P4d *)st, inc, obs, p,
                       (uint32_t)seed, (uint32_t)(seed >> 32), t, N);
}

void launch_plain_rest(FilterCtl *ctl, void *stv, double *lw, double *w, double *inc,
                       double *part, double *scal, uint64_t t, uint64_t N, int scheme,
                       ResampleBuffers &rb, StepRow *hist, hipStream_t s)
{
    P4d *st = (P4d *)stv;
    // log-evidence increment with the current normalised weights
    reduce<0>(inc, nullptr, nullptr, scal, kMaxInc, N, part, s);
    reduce<3>(inc, w, nullptr, scal, kSumW, N, part, s);
    // add_log_weight / set_log_weight, then normalise
    hipLaunchKernelGGL(k_plain_add, dim3(g256(N)), dim3(256), 0, s, lw, inc, N, t == 0 ? 1 : 0);
    reduce<0>(lw, nullptr, nullptr, scal, kMaxLw, N, part, s);
    hipLaunchKernelGGL(k_plain_sub, dim3(g256(N)), dim3(256), 0, s, lw, scal, N);
    hipLaunchKernelGGL(k_plain_exp, dim3(g256(N)), dim3(256), 0, s, lw, w, N);
    reduce<1>(w, nullptr, nullptr, scal, kAsum, N, part, s);
    hipLaunchKernelGGL(k_plain_scal, dim3(g256(N)), dim3(256), 0, s, w, scal, N);
    reduce<2>(w, w, nullptr, scal, kDot, N, part, s);
    hipLaunchKernelGGL(k_plain_decide, dim3(1), dim3(64), 0, s, ctl, scal, hist, t);
    // resample from the materialised weights, in-place copy, reset
    launch_resample_from_weights(ctl, w, N, scheme, rb, hist, s);
    hipLaunchKernelGGL(k_plain_copy, dim3(g256(N)), dim3(256), 0, s, ctl, st, rb.anc, N);
    hipLaunchKernelGGL(k_plain_reset, dim3(g256(N)), dim3(256), 0, s, ctl, w, lw, N);
    // monitor (MonitorMCMC)
    reduce<4>(nullptr, w, st, scal, kMx, N, part, s);
    hipLaunchKernelGGL(k_plain_mean, dim3(1), dim3(64), 0, s, scal, hist, t);
}

uint64_t plain_partials(uint64_t N) { return 2ull * gred(N); }

}  // namespace vsmcx
