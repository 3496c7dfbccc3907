// cv_tiles.hpp -- the per-tile and per-filter bodies of the CV step, shared by
// the single-filter kernels (kernels_cv.hip) and the serving tick
// (kernels_tick.hip), which run them on the same tiles in the same order.
//
// The move fuses, per particle, what vSMC spreads over MoveOMP's parallel
// loop and WeightSet's sequential passes: the ancestor gather of the previous
// resample (StateMatrix::copy, include/vsmc/core/state_matrix.hpp:285-300)
// folded into the state read, one Philox4x32-10 block per particle per step,
// the transition and Student-t log-likelihood of
// example/pf/include/pf_smp.hpp:96-112,207-238, lw += inc in FP64
// (core/weight_set.hpp:324-331), an online log-sum-exp (core/weight_set.hpp:
// 108-125), and the previous step's MonitorMCMC mean (core/monitor.hpp:274-302).
#pragma once
#include <type_traits>

#include "cv_model.hpp"
#include "device_util.hpp"
#include "internal.hpp"

namespace vsmcx {

// The source of each of the thread's particles in a move tile: with `gather`
// (uniform over the launch) the ancestor of the last resample, a zero slot's
// (kAncZero | its zero rank) taken from the surplus list at that rank
// (anc_source); else the particle itself.  All the thread's ancestor words are
// loaded before any is looked at, and the zero slots' list loads are again
// issued together.  Every result is an index into the state buffer (< N).
template <bool kFull = false>
__device__ __forceinline__ void tile_sources(const uint32_t *__restrict__ anc,
                                             const uint32_t *__restrict__ sl, bool gather,
                                             uint64_t N, uint64_t base, uint32_t src[kMoveItems])
{
    // kFull: the whole tile lies below N (every tile of a launch but the last)
    if (gather) {
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) {
            const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
            const uint32_t a = anc[kFull || i < N ? i : N - 1];  // in bounds, no branch
            src[k] = kFull || i < N ? a : (uint32_t)i;
        }
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) {
            const uint32_t a = src[k];
            if (a & kAncZero) {
                const uint32_t r = a & ~kAncZero;
                src[k] = r < N ? sl[r] : ~0u;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k)
            src[k] = (uint32_t)(base + (uint64_t)k * kBlock + threadIdx.x);
    }
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        const uint32_t j = src[k];
        // never read outside the state buffer
        src[k] = j < N ? j : (kFull || i < N ? (uint32_t)i : 0u);
    }
}

// One move tile (kMoveTile particles): writes the tile's partials
// (max, sum e, sum e^2, monitor x, monitor y) to part[0..4].  kFull: the
// tile lies below N, so no particle needs a bounds test.
template <typename Real, bool kFull>
__device__ __forceinline__ void cv_move_tile_body(
    FilterCtl *ctl, const typename StateOf<Real>::type *__restrict__ st_in,
    typename StateOf<Real>::type *__restrict__ st_out, const uint32_t *__restrict__ anc,
    const uint32_t *__restrict__ sl, double *__restrict__ lw, const double *__restrict__ obs,
    const CvParams &p, uint32_t k0, uint32_t k1, uint64_t t, uint64_t N, uint64_t tile,
    double *__restrict__ part)
{
    typedef typename StateOf<Real>::type S4;
    __shared__ double s_red[kBlock / kWave];
    __shared__ Part5 s_p5[kBlock / kWave];
    CvModel<Real> md;
    md.load(p);
    const int prev = t > 0 ? ctl->prev_resampled : 0;
    const double m_old = ctl->m, c_old = ctl->coeff;
    const LwView lw_old = lw_view(ctl, lw);  // the previous step's log-weights
    const double invN = 1.0 / (double)N;
    const CvObs<Real> o(obs[2 * t], obs[2 * t + 1]);
    const uint64_t base = tile * (uint64_t)kMoveTile;
    // With no old log-weights to add (first step, or after a resample) the FP32
    // pipeline's new log-weight is the increment alone, kept in compact form.
    float *const lwc = ctl->lwc;
    const bool compact = std::is_same<Real, float>::value && (t == 0 || prev) && lwc != nullptr;
    if (tile == 0 && threadIdx.x == 0) {  // read by the reduction after this launch
        ctl->lw_store_kind = compact ? 1 : 0;
        ctl->lw_c2 = md.c2;
        ctl->lw_h = md.half_nu1_ln2;
    }

    // Loads first, for all of the thread's particles (ancestor indices, then
    // the gathered states and old log-weights), so their latencies overlap;
    // then the arithmetic and the stores.
    uint32_t src[kMoveItems];
    tile_sources<kFull>(anc, sl, t > 0 && prev, N, base, src);
    // Every src[k] is in bounds, so the state loads need no branch (a load in
    // a branch of its own waits for the loads before it); past N the values
    // are not used.
    S4 s[kMoveItems];
    double lwo[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) lwo[k] = 0;
    if (t > 0) {  // uniform over the launch
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) s[k] = st_in[src[k]];
        if (!prev) {
#pragma unroll
            for (int k = 0; k < kMoveItems; ++k) {
                const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
                lwo[k] = lw_old(kFull || i < N ? i : N - 1);
            }
        }
    }

    double m = -INFINITY, mx = 0, my = 0;
    double lwn[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        lwn[k] = -INFINITY;
        if (!kFull && i >= N) continue;
        S4 &sk = s[k];
        if (t > 0) {
            const double w = prev ? invN : exp_lse(lwo[k] - m_old) * c_old;
            mx += w * (double)sk.x;
            my += w * (double)sk.y;
        }
        Real z[4];
        cv_normals(z, k0, k1, i, t);
        cv_advance(md, sk, z, t);
        st_out[i] = sk;
        if constexpr (std::is_same<Real, float>::value) {
            const float l2 = cv_llh_l2(md, sk.x, sk.y, o);
            lwn[k] = lwo[k] + (md.c2 - md.half_nu1_ln2 * (double)l2);
            if (compact)
                lwc[i] = l2;
            else
                lw[i] = lwn[k];
        } else {
            lwn[k] = lwo[k] + md.llh(sk.x, sk.y, o);
            lw[i] = lwn[k];
        }
        m = lwn[k] > m ? lwn[k] : m;
    }
    // the tile's maximum first, then one exp per particle relative to it
    const double M = block_reduce<kBlock>(m, MaxOp(), s_red);
    double se = 0, se2 = 0;
    if (M != -INFINITY) {
#pragma unroll
        for (int k = 0; k < kMoveItems; ++k) {
            const double e = exp_lse(lwn[k] - M);
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

template <typename Real>
__device__ __forceinline__ void cv_move_tile(
    FilterCtl *ctl, const typename StateOf<Real>::type *__restrict__ st_in,
    typename StateOf<Real>::type *__restrict__ st_out, const uint32_t *__restrict__ anc,
    const uint32_t *__restrict__ sl, double *__restrict__ lw, const double *__restrict__ obs,
    const CvParams &p, uint32_t k0, uint32_t k1, uint64_t t, uint64_t N, uint64_t tile,
    double *__restrict__ part)
{
    if ((tile + 1) * (uint64_t)kMoveTile <= N)
        cv_move_tile_body<Real, true>(ctl, st_in, st_out, anc, sl, lw, obs, p, k0, k1, t, N, tile,
                                      part);
    else
        cv_move_tile_body<Real, false>(ctl, st_in, st_out, anc, sl, lw, obs, p, k0, k1, t, N,
                                       tile, part);
}

// Tile partials {max, sum e, sum e^2, monitor x, monitor y} of ntiles tiles,
// combined in an order that the tile count alone fixes.  The tiles are dealt to
// nv = ceil(ntiles / 512) "virtual waves" (at most 16): lane l of virtual wave
// w takes tiles w*64 + l, then + 64*nv, ... in turn; the 64 lane results of a
// virtual wave meet in a fixed butterfly; the virtual waves' results are
// added in order of w.  Up to 512 tiles (a filter of up to 2^20 particles)
// that is a single wave: a 64-thread tick reduce, a 256-thread batch
// workgroup and the 1024-thread reduce of a single filter compute the same
// bits.  Larger filters spread their tiles over up to 16 waves of the
// single-filter reduce.
//
// Each lane loads kPartAhead of its tiles at once and evaluates their exps
// together before adding them in tile order: the order of the additions is
// that of the plain loop, but a wave does not wait a memory latency per tile.
constexpr int kPartAhead = 8;
constexpr int kPartMaxWaves = 16;
// the reduce after the move: one workgroup of kPartMaxWaves waves
constexpr int kReduceBlock = kPartMaxWaves * kWave;

__device__ __forceinline__ int part_vwaves(uint64_t ntiles)
{
    const uint64_t w = (ntiles + (uint64_t)kPartAhead * kWave - 1) / ((uint64_t)kPartAhead * kWave);
    return w < 1 ? 1 : (w > (uint64_t)kPartMaxWaves ? kPartMaxWaves : (int)w);
}

// The maximum over the tiles of virtual wave w of nv (every lane gets it).
// A: tiles loaded at once per lane; the additions below are in tile order
// whatever A is, so A only trades registers for memory latency.
template <int A = kPartAhead>
__device__ __forceinline__ double vwave_max(const double *partials, uint64_t ntiles, int w, int nv)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t S = (uint64_t)kWave * nv, kStep = (uint64_t)A * S;
    double m = -INFINITY;
    uint64_t j = (uint64_t)w * kWave + lane;
    for (; j + (A - 1) * S < ntiles; j += kStep) {
        double v[A];
#pragma unroll
        for (int u = 0; u < A; ++u) v[u] = partials[5 * (j + (uint64_t)u * S)];
#pragma unroll
        for (int u = 0; u < A; ++u) m = m < v[u] ? v[u] : m;
    }
    for (; j < ntiles; j += S) {
        const double v = partials[5 * j];
        m = m < v ? v : m;
    }
    return wave_reduce(m, MaxOp());
}

// The sums over the tiles of virtual wave w of nv, scaled to the maximum M.
template <int A = kPartAhead>
__device__ __forceinline__ Part5 vwave_sums(const double *partials, uint64_t ntiles, int w, int nv,
                                            double M)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t S = (uint64_t)kWave * nv, kStep = (uint64_t)A * S;
    Part5 a{0, 0, 0, 0};
    uint64_t j = (uint64_t)w * kWave + lane;
    for (; j + (A - 1) * S < ntiles; j += kStep) {
        double q[A][5], sc[A];
#pragma unroll
        for (int u = 0; u < A; ++u)
#pragma unroll
            for (int c = 0; c < 5; ++c) q[u][c] = partials[5 * (j + (uint64_t)u * S) + c];
#pragma unroll
        for (int u = 0; u < A; ++u) sc[u] = q[u][0] == -INFINITY ? 0.0 : exp_lse(q[u][0] - M);
#pragma unroll
        for (int u = 0; u < A; ++u) {
            a.se += q[u][1] * sc[u];
            a.se2 += q[u][2] * sc[u] * sc[u];
            a.mx += q[u][3];
            a.my += q[u][4];
        }
    }
    for (; j < ntiles; j += S) {
        const double *q = partials + 5 * j;
        const double sc = q[0] == -INFINITY ? 0.0 : exp_lse(q[0] - M);
        a.se += q[1] * sc;
        a.se2 += q[2] * sc * sc;
        a.mx += q[3];
        a.my += q[4];
    }
    return wave_reduce(a, Part5Add());
}

// The monitor fields of virtual wave w of nv.
template <int A = kPartAhead>
__device__ __forceinline__ Part5 vwave_monitor(const double *partials, uint64_t ntiles, int w, int nv)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t S = (uint64_t)kWave * nv, kStep = (uint64_t)A * S;
    Part5 a{0, 0, 0, 0};
    uint64_t j = (uint64_t)w * kWave + lane;
    for (; j + (A - 1) * S < ntiles; j += kStep) {
        double x[A], y[A];
#pragma unroll
        for (int u = 0; u < A; ++u) {
            x[u] = partials[5 * (j + (uint64_t)u * S) + 3];
            y[u] = partials[5 * (j + (uint64_t)u * S) + 4];
        }
#pragma unroll
        for (int u = 0; u < A; ++u) {
            a.mx += x[u];
            a.my += y[u];
        }
    }
    for (; j < ntiles; j += S) {
        a.mx += partials[5 * j + 3];
        a.my += partials[5 * j + 4];
    }
    return wave_reduce(a, Part5Add());
}

// One wave (all 64 lanes), ntiles <= 512: the single virtual wave.
__device__ __forceinline__ void tile_partials_lse(const double *partials, uint64_t ntiles,
                                                  double &M, Part5 &acc)
{
    M = vwave_max(partials, ntiles, 0, 1);
    acc = vwave_sums(partials, ntiles, 0, 1, M);
}

__device__ __forceinline__ void tile_partials_monitor(const double *partials, uint64_t ntiles,
                                                      double &mx, double &my)
{
    const Part5 a = vwave_monitor(partials, ntiles, 0, 1);
    mx = a.mx;
    my = a.my;
}

// LDS of the workgroup-wide forms
struct PartLds {
    double m[kPartMaxWaves];
    Part5 p[kPartMaxWaves];
};

// Every thread of a workgroup of any multiple of 64 threads; every thread
// gets the results.
__device__ __forceinline__ void tile_partials_lse_block(const double *partials, uint64_t ntiles,
                                                        double &M, Part5 &acc, PartLds &L)
{
    const int nv = part_vwaves(ntiles);
    const int pw = threadIdx.x / kWave, npw = blockDim.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = pw; w < nv; w += npw) {
        const double m = vwave_max<4>(partials, ntiles, w, nv);
        if (lane == 0) L.m[w] = m;
    }
    __syncthreads();
    M = L.m[0];
    for (int w = 1; w < nv; ++w) M = M < L.m[w] ? L.m[w] : M;
    for (int w = pw; w < nv; w += npw) {
        const Part5 a = vwave_sums<4>(partials, ntiles, w, nv, M);
        if (lane == 0) L.p[w] = a;
    }
    __syncthreads();
    acc = L.p[0];
    for (int w = 1; w < nv; ++w) acc = Part5Add()(acc, L.p[w]);
    __syncthreads();  // L may be reused
}

// tile_partials_lse_block when every lane holds at most kStashPer tiles (up
// to kStashTiles tiles, a filter of up to 2^24 particles): each lane loads all
// of its partials in one round, takes the same maxima and the same additions
// in the same order (the same bits), and leaves each tile's scaled sum
// se_t e^(M_t - M) in stash[t], the product the fold's tile starts multiply
// by 1 / sum (TileSumsFromLse).  blockDim.x >= 64 * part_vwaves(ntiles).
constexpr int kStashPer = 8;
constexpr int kStashTiles = kStashPer * kWave * kPartMaxWaves;
__device__ __forceinline__ void tile_partials_lse_block_stash(const double *partials,
                                                              uint64_t ntiles, double &M,
                                                              Part5 &acc, PartLds &L,
                                                              double *stash)
{
    const int nv = part_vwaves(ntiles);
    const int pw = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const uint64_t S = (uint64_t)kWave * nv, j0 = (uint64_t)pw * kWave + lane;
    const bool act = pw < nv;
    double q[kStashPer][5];
    if (act) {
#pragma unroll
        for (int u = 0; u < kStashPer; ++u) {
            const uint64_t j = j0 + (uint64_t)u * S;
            const uint64_t jj = j < ntiles ? j : ntiles - 1;  // in bounds, no branch
#pragma unroll
            for (int c = 0; c < 5; ++c) q[u][c] = partials[5 * jj + c];
        }
        double m = -INFINITY;
#pragma unroll
        for (int u = 0; u < kStashPer; ++u)
            if (j0 + (uint64_t)u * S < ntiles) m = m < q[u][0] ? q[u][0] : m;
        m = wave_reduce(m, MaxOp());
        if (lane == 0) L.m[pw] = m;
    }
    __syncthreads();
    M = L.m[0];
    for (int w = 1; w < nv; ++w) M = M < L.m[w] ? L.m[w] : M;
    if (act) {
        Part5 a{0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < kStashPer; ++u) {
            const uint64_t j = j0 + (uint64_t)u * S;
            if (j < ntiles) {
                const double sc = q[u][0] == -INFINITY ? 0.0 : exp_lse(q[u][0] - M);
                const double s1 = q[u][1] * sc;
                a.se += s1;
                a.se2 += q[u][2] * sc * sc;
                a.mx += q[u][3];
                a.my += q[u][4];
                stash[j] = s1;
            }
        }
        a = wave_reduce(a, Part5Add());
        if (lane == 0) L.p[pw] = a;
    }
    __syncthreads();
    acc = L.p[0];
    for (int w = 1; w < nv; ++w) acc = Part5Add()(acc, L.p[w]);
    __syncthreads();  // L may be reused
}

__device__ __forceinline__ void tile_partials_monitor_block(const double *partials,
                                                            uint64_t ntiles, double &mx,
                                                            double &my, PartLds &L)
{
    const int nv = part_vwaves(ntiles);
    const int pw = threadIdx.x / kWave, npw = blockDim.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    for (int w = pw; w < nv; w += npw) {
        const Part5 a = vwave_monitor<4>(partials, ntiles, w, nv);
        if (lane == 0) L.p[w] = a;
    }
    __syncthreads();
    Part5 acc = L.p[0];
    for (int w = 1; w < nv; ++w) acc = Part5Add()(acc, L.p[w]);
    mx = acc.mx;
    my = acc.my;
    __syncthreads();
}

// After a step's move: ESS, log-evidence and the resample decision from the
// tile partials (wave 0 of the calling workgroup; thread 0 writes); up to
// 512 tiles.
__device__ __forceinline__ void cv_reduce_tiles(FilterCtl *ctl, const double *partials,
                                                uint64_t ntiles, StepRow *hist, uint64_t t)
{
    if (threadIdx.x >= kWave) return;
    double M;
    Part5 acc;
    tile_partials_lse(partials, ntiles, M, acc);
    if (threadIdx.x == 0) cv_step_scalars(ctl, M, acc.se, acc.se2, acc.mx, acc.my, hist, t);
}

// Monitor of the last step: sum_i W_i (x_i, y_i) over the post-resample state.
template <typename Real>
__device__ __forceinline__ void cv_monitor_tile(const FilterCtl *ctl,
                                                const typename StateOf<Real>::type *st,
                                                const uint32_t *anc, const uint32_t *sl,
                                                const double *lw, uint64_t N, uint64_t tile,
                                                double *part)
{
    __shared__ Part5 s_p5[kBlock / kWave];
    const int prev = ctl->prev_resampled;
    const double m = ctl->m, c = ctl->coeff, invN = 1.0 / (double)N;
    const LwView lwf = lw_view(ctl, lw);
    const uint64_t base = tile * (uint64_t)kMoveTile;
    Part5 v{0, 0, 0, 0};
    // loads first (sources, states, log-weights), then the sums in particle order
    uint32_t src[kMoveItems];
    tile_sources(anc, sl, prev != 0, N, base, src);
    typename StateOf<Real>::type s[kMoveItems];
    double lwv[kMoveItems];
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        s[k] = st[src[k]];
        lwv[k] = prev ? 0.0 : lwf(i < N ? i : N - 1);
    }
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i >= N) continue;
        const double w = prev ? invN : exp_lse(lwv[k] - m) * c;
        v.mx += w * (double)s[k].x;
        v.my += w * (double)s[k].y;
    }
    v = block_reduce<kBlock>(v, Part5Add(), s_p5);
    if (threadIdx.x == 0) {
        part[3] = v.mx;
        part[4] = v.my;
    }
}

__device__ __forceinline__ void cv_monitor_reduce(const FilterCtl *ctl, const double *partials,
                                                  uint64_t ntiles, StepRow *hist)
{
    if (threadIdx.x >= kWave) return;
    double mx, my;
    tile_partials_monitor(partials, ntiles, mx, my);
    if (threadIdx.x == 0 && hist) {
        hist[ctl->t].mean_x = mx;
        hist[ctl->t].mean_y = my;
    }
}

}  // namespace vsmcx
