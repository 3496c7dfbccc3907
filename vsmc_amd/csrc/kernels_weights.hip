// kernels_weights.hip -- the device weight set (vsmcx_weights_*): log-weights
// that a user kernel's increments update, with the filter's own reduction
// and resampler behind them, and the ancestor gather of a state in place.
#include "cv_tiles.hpp"
#include "launch.hpp"

namespace vsmcx {

// WeightSet::set_log_weight (set != 0) or add_log_weight
// (include/vsmc/core/weight_set.hpp:297-331) over one 2048-particle tile:
// lw_i = (set, or the last step resampled ? 0 : lw_i) + inc_i, where a
// resample left the weights equal (set_equal_weight), and the tile's
// partials {max, sum exp(lw - max), sum exp(2 (lw - max)), 0, 0} in the move
// kernel's layout: the maximum first, then one exp per particle.
__global__ __launch_bounds__(kBlock) void k_ws_add(const FilterCtl *ctl, double *lw,
                                                   const double *inc, uint64_t N, int set,
                                                   double *part)
{
    __shared__ double s_red[kBlock / kWave];
    __shared__ Part5 s_p5[kBlock / kWave];
    const int reset = set || ctl->prev_resampled;
    const uint64_t base = blockIdx.x * (uint64_t)kMoveTile;
    double v[kMoveItems];
    double m = -INFINITY;
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        v[k] = -INFINITY;
        if (i >= N) continue;
        v[k] = (reset ? 0.0 : lw[i]) + inc[i];
        lw[i] = v[k];
        m = v[k] > m ? v[k] : m;
    }
    const double M = block_reduce<kBlock>(m, MaxOp(), s_red);
    double se = 0, se2 = 0;
    if (M != -INFINITY) {
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) {
            const double e = exp_lse(v[k] - M);
            se += e;
            se2 += e * e;
        }
    }
    Part5 p{se, se2, 0, 0};
    p = block_reduce<kBlock>(p, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        double *q = part + 5 * (uint64_t)blockIdx.x;
        q[0] = M;
        q[1] = p.se;
        q[2] = p.se2;
        q[3] = 0;
        q[4] = 0;
    }
}

// StateMatrix::copy (include/vsmc/core/state_matrix.hpp:285-300) in place
// when this step resampled: x[to] = x[anc[to]] for anc[to] != to.  The
// cfrp_trans layout keeps every source in place (a particle with copies
// keeps itself, anc[s] = s), so no row is both read and rewritten, and
// threads need no order.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_ws_gather(const FilterCtl *ctl, Real *x,
                                                      const uint32_t *anc, uint32_t dim,
                                                      uint64_t N, int *err)
{
    if (!ctl->resample) return;
    for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < N * dim;
         e += (uint64_t)gridDim.x * kBlock) {
        const uint64_t to = e / dim, d = e - to * dim;
        const uint64_t from = anc[to];
        if (from >= N)
            atomicOr(err, (int)kErrBounds);
        else if (from != to)
            x[e] = x[from * dim + d];
    }
}

// the step's resampling is done: a second resample call before the next
// add_log_weight does nothing
__global__ void k_ws_done(FilterCtl *ctl)
{
    if (threadIdx.x == 0) ctl->resample = 0;
}

void launch_ws_add(const FilterCtl *ctl, double *lw, const double *inc, uint64_t N, int set,
                   double *part, hipStream_t s)
{
    const unsigned tiles = (unsigned)((N + kMoveTile - 1) / kMoveTile);
    hipLaunchKernelGGL(k_ws_add, dim3(tiles), dim3(kBlock), 0, s, ctl, lw, inc, N, set, part);
}

void launch_ws_gather(bool fp64, const FilterCtl *ctl, void *x, const uint32_t *anc,
                      uint32_t dim, uint64_t N, int *err, hipStream_t s)
{
    const uint64_t n = (N * dim + kBlock - 1) / kBlock;
    const unsigned grid = (unsigned)(n < (1u << 16) ? n : (1u << 16));
    if (fp64)
        hipLaunchKernelGGL(k_ws_gather<double>, dim3(grid ? grid : 1), dim3(kBlock), 0, s, ctl,
                           (double *)x, anc, dim, N, err);
    else
        hipLaunchKernelGGL(k_ws_gather<float>, dim3(grid ? grid : 1), dim3(kBlock), 0, s, ctl,
                           (float *)x, anc, dim, N, err);
}

void launch_ws_done(FilterCtl *ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_ws_done, dim3(1), dim3(64), 0, s, ctl);
}

}  // namespace vsmcx
