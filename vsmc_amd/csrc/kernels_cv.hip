// kernels_cv.hip -- the single-filter kernels of the constant-velocity model
// (example/pf): move + likelihood + online log-sum-exp per tile, the one-
// workgroup reduction with the resample decision, the final monitor, and the
// helpers of the debug and state entry points.  The bodies are in
// cv_tiles.hpp.  State is AoS: 4 x FP32 (16 B) in the mixed pipeline, 4 x
// FP64 (32 B) in the FP64 pipelines.
#include <algorithm>
#include <math.h>

#include "device_util.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "cv_model.hpp"
#include "cv_tiles.hpp"
#include "fold_tiles.hpp"
#include "rng_core.hpp"

namespace vsmcx {

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_cv_move(
    FilterCtl *ctl, const typename StateOf<Real>::type *__restrict__ st_in,
    typename StateOf<Real>::type *__restrict__ st_out, const uint32_t *__restrict__ anc,
    const uint32_t *__restrict__ sl, double *__restrict__ lw, const double *__restrict__ obs,
    CvParams p, uint32_t k0, uint32_t k1, uint64_t t, uint64_t N, double *__restrict__ partials)
{
    cv_move_tile<Real>(ctl, st_in, st_out, anc, sl, lw, obs, p, k0, k1, t, N, blockIdx.x,
                       partials + 5 * (uint64_t)blockIdx.x);
}

// After the move: ESS, log-evidence and the resample decision from the tile
// partials; with prologue_scheme >= 0 and a resample, also the scheme's
// uniform-grid prologue and the fold's approximate tile starts of the
// filter's weights (TileSumsFromLse) into fs.ts, which the fold would
// otherwise compute in two more launches.
__global__ __launch_bounds__(kReduceBlock) void k_cv_reduce(FilterCtl *ctl, const double *partials,
                                                         uint64_t ntiles, StepRow *hist,
                                                         uint64_t t, int prologue_scheme,
                                                         FoldScratch fs)
{
    __shared__ PartLds L;
    __shared__ double s_red[kReduceBlock / kWave];
    __shared__ double s_mc[2];
    __shared__ int s_res;
    __shared__ double s_stash[kStashTiles];  // se_t e^(M_t - M) per tile
    const bool stash = ntiles <= (uint64_t)kStashTiles;
    double M;
    Part5 acc;
    if (stash)
        tile_partials_lse_block_stash(partials, ntiles, M, acc, L, s_stash);
    else
        tile_partials_lse_block(partials, ntiles, M, acc, L);
    if (threadIdx.x == 0) {
        cv_step_scalars(ctl, M, acc.se, acc.se2, acc.mx, acc.my, hist, t);
        const int res = ctl->resample;
        if (res && prologue_scheme >= 0) resample_prologue(ctl, prologue_scheme);
        s_mc[0] = ctl->m;
        s_mc[1] = ctl->coeff;
        s_res = res;
    }
    if (prologue_scheme < 0 || !fs.ts) return;
    __syncthreads();
    if (!s_res) return;
    if (stash) {
        const TileSumsStashed in{s_stash, s_mc[1]};
        wg_scan<double>(in, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
    } else {
        const TileSumsFromLse in{partials, s_mc[0], s_mc[1]};
        wg_scan<double>(in, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
    }
}

// Monitor of the last step: sum_i W_i (x_i, y_i) over the post-resample state.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_cv_monitor(FilterCtl *ctl,
                                                       const typename StateOf<Real>::type *st,
                                                       const uint32_t *anc, const uint32_t *sl,
                                                       const double *lw, uint64_t N,
                                                       double *partials)
{
    cv_monitor_tile<Real>(ctl, st, anc, sl, lw, N, blockIdx.x,
                          partials + 5 * (uint64_t)blockIdx.x);
}

__global__ __launch_bounds__(kReduceBlock) void k_cv_monitor_sum(FilterCtl *ctl,
                                                              const double *partials,
                                                              uint64_t ntiles, StepRow *hist)
{
    __shared__ PartLds L;
    double mx, my;
    tile_partials_monitor_block(partials, ntiles, mx, my, L);
    if (threadIdx.x == 0 && hist) {
        hist[ctl->t].mean_x = mx;
        hist[ctl->t].mean_y = my;
    }
}

// W_i exactly as the resampler computes it (XWeight in kernels_fold.hip)
__global__ void k_cv_weights(const FilterCtl *ctl, const double *lw, uint64_t N, double *w)
{
    const double m = ctl->m, c = ctl->coeff;
    const LwView lwv = lw_view(ctl, lw);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = exp_lse(lwv(i) - m) * c;
}

// The log-weights in FP64 whatever their form (exports and checkpoints).
__global__ void k_lw_expand(const FilterCtl *ctl, const double *lw, uint64_t N, double *out)
{
    const LwView lwv = lw_view(ctl, lw);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = lwv(i);
}

template <typename S4>
__global__ void k_cv_gather(const FilterCtl *ctl, const S4 *in, S4 *out, const uint32_t *anc,
                            const uint32_t *sl, uint64_t N)
{
    const int prev = ctl->prev_resampled;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = in[prev ? (uint64_t)anc_source(anc, sl, i) : i];
}

static unsigned tiles_of(uint64_t N) { return (unsigned)((N + kMoveTile - 1) / kMoveTile); }

void launch_cv_move(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                    const uint32_t *anc, const uint32_t *sl, double *lw, const double *obs,
                    const CvParams &p, uint64_t seed, uint64_t t, uint64_t N, double *partials,
                    hipStream_t st)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (fp64)
        hipLaunchKernelGGL(k_cv_move<double>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4d *)st_in, (S4d *)st_out, anc, sl, lw, obs, p, k0, k1, t, N,
                           partials);
    else
        hipLaunchKernelGGL(k_cv_move<float>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4f *)st_in, (S4f *)st_out, anc, sl, lw, obs, p, k0, k1, t, N,
                           partials);
}

void launch_cv_reduce(FilterCtl *ctl, const double *partials, uint64_t ntiles, StepRow *hist,
                      uint64_t t, hipStream_t st, int prologue_scheme, const FoldScratch *starts)
{
    FoldScratch fs{};
    if (starts) fs = *starts;
    hipLaunchKernelGGL(k_cv_reduce, dim3(1), dim3(kReduceBlock), 0, st, ctl, partials, ntiles, hist,
                       t, starts ? prologue_scheme : -1, fs);
}

void launch_cv_monitor(bool fp64, FilterCtl *ctl, const void *stp, const uint32_t *anc,
                       const uint32_t *sl, const double *lw, uint64_t N, double *partials,
                       StepRow *hist, hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_cv_monitor<double>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4d *)stp, anc, sl, lw, N, partials);
    else
        hipLaunchKernelGGL(k_cv_monitor<float>, dim3(tiles_of(N)), dim3(kBlock), 0, st, ctl,
                           (const S4f *)stp, anc, sl, lw, N, partials);
    hipLaunchKernelGGL(k_cv_monitor_sum, dim3(1), dim3(kReduceBlock), 0, st, ctl, partials,
                       (uint64_t)tiles_of(N), hist);
}

void launch_lw_expand(const FilterCtl *ctl, const double *lw, uint64_t N, double *out,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_lw_expand, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl, lw, N, out);
}

void launch_cv_weights(FilterCtl *ctl, const double *lw, uint64_t N, double *w, hipStream_t st)
{
    hipLaunchKernelGGL(k_cv_weights, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl, lw, N, w);
}

void launch_cv_gather(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                      const uint32_t *anc, const uint32_t *sl, uint64_t N, hipStream_t st)
{
    if (fp64)
        hipLaunchKernelGGL(k_cv_gather<S4d>, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl,
                           (const S4d *)st_in, (S4d *)st_out, anc, sl, N);
    else
        hipLaunchKernelGGL(k_cv_gather<S4f>, dim3(tiles_of(N) * 16), dim3(256), 0, st, ctl,
                           (const S4f *)st_in, (S4f *)st_out, anc, sl, N);
}

// ---- device-resident particle values (the vSMC value-type adaptor) ------------------
// One pf-model step over a bare state (no weight set on the device): the
// transition (or the prior draw at t == 0) and the log-likelihood increment of
// every particle, with the move kernel's normals and arithmetic.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_state_cv_move(typename StateOf<Real>::type *st,
                                                          double *inc, CvParams p, uint32_t k0,
                                                          uint32_t k1, uint64_t t, double ox,
                                                          double oy, uint64_t N)
{
    CvModel<Real> md;
    md.load(p);
    const CvObs<Real> o(ox, oy);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * kBlock) {
        typename StateOf<Real>::type s = st[i];
        Real z[4];
        cv_normals(z, k0, k1, i, t);
        cv_advance(md, s, z, t);
        st[i] = s;
        inc[i] = md.llh(s.x, s.y, o);
    }
}

// StateMatrix::copy (core/state_matrix.hpp:285-300) as a gather into the
// other buffer: out[to] = in[copy_from[to]], rows of dim values.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_state_gather(const Real *in, Real *out,
                                                         const uint64_t *copy_from, uint32_t dim,
                                                         uint64_t N, int *err)
{
    for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < N * dim;
         e += (uint64_t)gridDim.x * kBlock) {
        const uint64_t to = e / dim, d = e - to * dim;
        const uint64_t from = copy_from[to];
        if (from < N)
            out[e] = in[from * dim + d];
        else
            atomicOr(err, (int)kErrBounds);
    }
}

static unsigned grid_of(uint64_t n) { return (unsigned)std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 16); }

void launch_state_cv_move(bool fp64, void *st, double *inc, const CvParams &p, uint64_t seed,
                          uint64_t t, double ox, double oy, uint64_t N, hipStream_t s)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (fp64)
        hipLaunchKernelGGL(k_state_cv_move<double>, dim3(grid_of(N)), dim3(kBlock), 0, s,
                           (S4d *)st, inc, p, k0, k1, t, ox, oy, N);
    else
        hipLaunchKernelGGL(k_state_cv_move<float>, dim3(grid_of(N)), dim3(kBlock), 0, s,
                           (S4f *)st, inc, p, k0, k1, t, ox, oy, N);
}

void launch_state_gather(bool fp64, const void *in, void *out, const uint64_t *copy_from,
                         uint32_t dim, uint64_t N, int *err, hipStream_t s)
{
    if (fp64)
        hipLaunchKernelGGL(k_state_gather<double>, dim3(grid_of(N * dim)), dim3(kBlock), 0, s,
                           (const double *)in, (double *)out, copy_from, dim, N, err);
    else
        hipLaunchKernelGGL(k_state_gather<float>, dim3(grid_of(N * dim)), dim3(kBlock), 0, s,
                           (const float *)in, (float *)out, copy_from, dim, N, err);
}

}  // namespace vsmcx
