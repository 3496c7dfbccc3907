// lg_tiles.hpp -- the per-tile bodies of the linear-Gaussian model's step
// (x_t = A x_{t-1} + q z, y_t = x_t + r v), shared by the serving tick
// (kernels_tick.hip) and the persistent batch kernel (kernels_batch.hip),
// which run them on the same tiles in the same order; the tile partials
// have the layout of the CV tiles (cv_tiles.hpp).
#pragma once
#include "cv_model.hpp"
#include "cv_tiles.hpp"  // tile_sources
#include "device_util.hpp"
#include "launch.hpp"
#include "lg_model.hpp"

namespace vsmcx {

// ---- the linear-Gaussian move on one tile (same tile shape as cv_move_tile) ----
template <typename Real>
__device__ __forceinline__ void lg_move_tile(const FilterCtl *ctl, const Real *__restrict__ in,
                                             Real *__restrict__ out, const uint32_t *__restrict__ anc,
                                             const uint32_t *__restrict__ sl,
                                             double *__restrict__ lw, const double *__restrict__ y,
                                             const LgParams &lp, uint32_t k0, uint32_t k1,
                                             uint64_t t, uint64_t N, uint64_t tile, double *part)
{
    __shared__ double s_red[kBlock / kWave];
    __shared__ Part5 s_p5[kBlock / kWave];
    const int D = lp.dim;
    const int prev = t > 0 ? ctl->prev_resampled : 0;
    const double m_old = ctl->m, c_old = ctl->coeff;
    const double invN = 1.0 / (double)N;
    const Real q = (Real)lp.q, s0 = (Real)lp.s0;
    const double inv_r2 = 1 / (lp.r * lp.r);
    const uint64_t base = tile * (uint64_t)kMoveTile;
    double m = -INFINITY, mx = 0, my = 0;
    double lwv[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) lwv[k] = -INFINITY;
#pragma unroll 1
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i >= N) continue;
        Real x[kLgMaxDim], xn[kLgMaxDim];
        double lwo = 0;
        if (t > 0) {
            uint64_t src = prev ? (uint64_t)anc_source(anc, sl, i) : i;
            src = src < N ? src : i;
            // the state's loads at dimensions clamped into [0, D): no branch
            // between them
#pragma unroll
            for (int d = 0; d < kLgMaxDim; ++d) {
                const Real xv = in[src * D + (d < D ? d : D - 1)];
                x[d] = d < D ? xv : (Real)0;
            }
            if (!prev) lwo = lw[i];
            const double w = prev ? invN : exp_lse(lwo - m_old) * c_old;
            mx += w * (double)x[0];
            my += w * (double)x[1];
        }
        Real z[12];
#pragma unroll
        for (int j = 0; j < 3; ++j) lg_normals(z + 4 * j, k0, k1, i, t, (uint32_t)j);
        Real ss = 0;
#pragma unroll
        for (int d = 0; d < kLgMaxDim; ++d) {
            if (d >= D) break;
            Real v;
            if (t == 0) {
                v = s0 * z[d];
            } else {
                v = q * z[d];
#pragma unroll
                for (int e = 0; e < kLgMaxDim; ++e)
                    if (e < D) v += lg_a<Real>(lp, d * D + e) * x[e];
            }
            xn[d] = v;
            // y_d as hi + lo: y_hi - v is exact near the observation, so the
            // observation's own FP32 rounding, shared by all particles, does
            // not enter the residual (lo = 0 in FP64)
            const double yd = y[d];
            const Real yh = (Real)yd, yl = (Real)(yd - (double)yh);
            const Real r = (yh - v) + yl;
            ss += r * r;
        }
        const double lwn = lwo + (lp.llh_const - 0.5 * inv_r2 * (double)ss);
#pragma unroll
        for (int d = 0; d < kLgMaxDim; ++d)
            if (d < D) out[i * D + d] = xn[d];
        lw[i] = lwn;
        lwv[k] = lwn;
        m = lwn > m ? lwn : m;
    }
    // the tile's maximum first, then one exp per particle relative to it
    const double M = block_reduce<kBlock>(m, MaxOp(), s_red);
    double se = 0, se2 = 0;
    if (M != -INFINITY) {
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) {
            const double e = exp_lse(lwv[k] - M);
            se += e;
            se2 += e * e;
        }
    }
    Part5 v{se, se2, mx, my};
    v = block_reduce<kBlock>(v, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        part[0] = M;
        part[1] = v.se;
        part[2] = v.se2;
        part[3] = v.mx;
        part[4] = v.my;
    }
}

// final monitor of a linear-Gaussian job: mean of the first two coordinates
template <typename Real>
__device__ __forceinline__ void lg_monitor_tile(const FilterCtl *ctl, const Real *st, int D,
                                                const uint32_t *anc, const uint32_t *sl,
                                                const double *lw, uint64_t N, uint64_t tile,
                                                double *part)
{
    __shared__ Part5 s_p5[kBlock / kWave];
    const int prev = ctl->prev_resampled;
    const double m = ctl->m, c = ctl->coeff, invN = 1.0 / (double)N;
    const uint64_t base = tile * (uint64_t)kMoveTile;
    Part5 v{0, 0, 0, 0};
    // loads first (sources, the first two coordinates, log-weights), then the
    // sums in particle order; the second coordinate is 0 for a one-dimensional
    // state, as in the move
    uint32_t src[kMoveItems];
    tile_sources(anc, sl, prev != 0, N, base, src);
    Real x0[kMoveItems], x1[kMoveItems];
    double lwv[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        x0[k] = st[(uint64_t)src[k] * D];
        x1[k] = st[(uint64_t)src[k] * D + (D > 1 ? 1 : 0)];
        lwv[k] = prev ? 0.0 : lw[i < N ? i : N - 1];
    }
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i >= N) continue;
        const double w = prev ? invN : exp_lse(lwv[k] - m) * c;
        v.mx += w * (double)x0[k];
        v.my += w * (D > 1 ? (double)x1[k] : 0.0);
    }
    v = block_reduce<kBlock>(v, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        part[3] = v.mx;
        part[4] = v.my;
    }
}

}  // namespace vsmcx
