// kernels_batch.hip -- thousands of independent filters in one persistent launch.
//
// Each workgroup takes filters from a queue (largest first, or the order the
// tuner chose) and runs whole steps of one filter at a time: the move and
// likelihood of every particle, the FP64 log-sum-exp and ESS, and, when the
// ESS falls below the threshold, the resample -- the same per-tile fold,
// inversion and cfrp_trans bodies the large-filter grid passes run
// (fold_tiles.hpp), here tile after tile inside the workgroup, with the
// tile-level scans in LDS.  Every reduction has a fixed order that depends on
// the filter alone, so a filter's results are the same whichever workgroup
// runs it, in whatever company, under whatever launch shape.
#include "cv_tiles.hpp"
#include "fold_tiles.hpp"
#include "launch.hpp"
#include "lg_tiles.hpp"

namespace vsmcx {

template <typename Inv>
struct BatchLds {
    FoldLdsT<Inv::kWindowTmp> L;
    SpecRec r[kBlock];
    double v[kBlock];
    Part5 p5[kBlock / kWave];
    Seg tail;
    Seg carry;
    uint64_t zeros;
    unsigned job;
    int flag;
};

template <typename Inv>
struct SchemeOf;
template <>
struct SchemeOf<InvSystematic> {
    static constexpr int value = kSystematic;
};
template <>
struct SchemeOf<InvStratified> {
    static constexpr int value = kStratified;
};

// Sum over the current state of W_i (x_i, y_i) -- MonitorMCMC's mean after the
// last step (core/monitor.hpp:274-302); x, y are the first two coordinates.
// The single-filter and serving monitor tiles, tile after tile, then their
// fixed-order reduction.
template <typename Real, typename Inv>
__device__ __noinline__ void batch_final_monitor(const BatchFilter &F, uint32_t f,
                                                 const FilterCtl *ctl, const BatchPool &P,
                                                 const LgParams &lp, StepRow *row,
                                                 BatchLds<Inv> &S)
{
    const uint64_t N = F.n, T = (N + kMoveTile - 1) / kMoveTile;
    const Real *st = (P.cur[f] ? (const Real *)P.st1 : (const Real *)P.st0) + F.soff;
    double *part = P.tpart + 5 * (F.off / kMoveTile);
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt) {
        if (F.model == kModelLg)
            lg_monitor_tile<Real>(ctl, st, lp.dim, P.anc + F.off, P.sl + F.off, P.lw + F.off, N, tt,
                                  part + 5 * tt);
        else
            cv_monitor_tile<Real>(ctl, (const typename StateOf<Real>::type *)st, P.anc + F.off,
                                  P.sl + F.off, P.lw + F.off, N, tt, part + 5 * tt);
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
        double mx, my;
        tile_partials_monitor(part, T, mx, my);
        if (threadIdx.x == 0) {
            row->mean_x = mx;
            row->mean_y = my;
        }
    }
    __syncthreads();
}

// The reductions that end every move: the tile partials in their fixed order
// (thread 0 writes the step's scalars).
template <typename Inv>
__device__ __forceinline__ void batch_move_end(FilterCtl *ctl, const BatchPool &P, uint32_t f,
                                               uint32_t cur, StepRow *hist, uint64_t t,
                                               const double *part, uint64_t ntiles,
                                               BatchLds<Inv> &S)
{
    __syncthreads();  // every tile's partials written
    if (threadIdx.x < kWave) {
        double M;
        Part5 v;
        tile_partials_lse(part, ntiles, M, v);
        if (threadIdx.x == 0) {
            cv_step_scalars(ctl, M, v.se, v.se2, v.mx, v.my, hist, t);
            P.cur[f] = 1 - cur;
            S.flag = ctl->resample;
            if (ctl->resample) resample_prologue(ctl, SchemeOf<Inv>::value);
        }
    }
    __syncthreads();
}

// The phases of a step are separate (not inlined) functions: each gets its own
// register allocation, and what crosses phases lives in LDS, so the kernel
// keeps the occupancy of its largest phase instead of the sum of all.

// CV model: the single filter's move tiles, tile after tile, then the
// decision.
template <typename Real, typename Inv>
__device__ __noinline__ void batch_move_cv(const BatchFilter &F, uint32_t f, const BatchPool &P,
                                           const CvParams &p, uint64_t t, BatchLds<Inv> &S)
{
    typedef typename StateOf<Real>::type S4;
    FilterCtl *ctl = P.ctl + f;
    const uint64_t N = F.n, T = (N + kMoveTile - 1) / kMoveTile;
    S4 *sa = (S4 *)((Real *)P.st0 + F.soff), *sb = (S4 *)((Real *)P.st1 + F.soff);
    const uint32_t cur = P.cur[f];
    const uint32_t k0 = (uint32_t)F.seed, k1 = (uint32_t)(F.seed >> 32);
    double *part = P.tpart + 5 * (F.off / kMoveTile);
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        cv_move_tile<Real>(ctl, cur ? sb : sa, cur ? sa : sb, P.anc + F.off, P.sl + F.off, P.lw + F.off,
                           P.obs + F.obs_off, p, k0, k1, t, N, tt, part + 5 * tt);
    batch_move_end(ctl, P, f, cur, P.hist + F.hist_off, t, part, T, S);
}

// Linear-Gaussian model: the serving tick's move tiles, tile after tile.
template <typename Real, typename Inv>
__device__ __noinline__ void batch_move_lg(const BatchFilter &F, uint32_t f, const BatchPool &P,
                                           const LgParams &lp, uint64_t t, BatchLds<Inv> &S)
{
    FilterCtl *ctl = P.ctl + f;
    const uint64_t N = F.n, T = (N + kMoveTile - 1) / kMoveTile;
    Real *sa = (Real *)P.st0 + F.soff, *sb = (Real *)P.st1 + F.soff;
    const uint32_t cur = P.cur[f];
    const uint32_t k0 = (uint32_t)F.seed, k1 = (uint32_t)(F.seed >> 32);
    double *part = P.tpart + 5 * (F.off / kMoveTile);
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        lg_move_tile<Real>(ctl, cur ? sb : sa, cur ? sa : sb, P.anc + F.off, P.sl + F.off, P.lw + F.off,
                           P.obs + F.obs_off + (uint64_t)lp.dim * t, lp, k0, k1, t, N, tt,
                           part + 5 * tt);
    batch_move_end(ctl, P, f, cur, P.hist + F.hist_off, t, part, T, S);
}

// approximate tile starts: the exclusive scan of the tile sums, which follow
// from the move's tile partials (TileSumsFromLse)
template <typename Inv>
__device__ __noinline__ void batch_fold_sums(const BatchFilter &F, FilterCtl *ctl,
                                             const BatchPool &P, BatchLds<Inv> &S)
{
    const int tid = threadIdx.x;
    const uint64_t N = F.n, T = (N + kFT - 1) / kFT;
    const TileSumsFromLse in{P.tpart + 5 * (F.off / kFT), ctl->m, ctl->coeff};
    double *ts = P.ts + F.off / kFT;
    // exclusive scan over the tiles, kBlock at a time, in a fixed order
    double carry = 0.0;
#pragma unroll 1
    for (uint64_t c0 = 0; c0 < T; c0 += kBlock) {
        const uint64_t i = c0 + tid;
        const double a = i < T ? in[i] : 0.0;
        double tot;
        const double ex = block_excl_scan<kBlock>(a, AddOp(), 0.0, S.L.red, &tot);
        if (i < T) ts[i] = carry + ex;
        carry += tot;
    }
    __syncthreads();
}

// classification and segment summaries; their exclusive scan
template <typename Inv>
__device__ __noinline__ void batch_fold_segs(const BatchFilter &F, FilterCtl *ctl,
                                             const BatchPool &P, BatchLds<Inv> &S)
{
    const int tid = threadIdx.x;
    const uint64_t N = F.n, T = (N + kFT - 1) / kFT;
    XWeight xs{P.lw + F.off, 0.0, 0.0};
    xs.bind(ctl);
    const double *ts = P.ts + F.off / kFT;
    Seg *agg = P.tagg + F.off / kFT, *pre = P.tpre + F.off / kFT;
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        fold_tile_seg(xs, tt * (uint64_t)kFT, N, ts[tt], P.spec + F.off + tt * (uint64_t)kFT,
                      agg + tt, S.L);
    __syncthreads();
    if (tid == 0) S.carry = seg_identity();
    __syncthreads();
#pragma unroll 1
    for (uint64_t c0 = 0; c0 < T; c0 += kBlock) {
        const uint64_t i = c0 + tid;
        const Seg a = i < T ? agg[i] : seg_identity();
        const Seg ex = block_excl_scan<kBlock>(a, SegOp(), seg_identity(), S.L.seg, (Seg *)nullptr);
        const Seg c = S.carry;
        if (i < T) pre[i] = seg_op(c, ex);
        __syncthreads();
        if (tid == kBlock - 1) S.carry = seg_op(c, seg_op(ex, a));
        __syncthreads();
    }
    if (tid == 0) S.tail = S.carry;
    __syncthreads();
}

// the chain over the specials
template <typename Inv>
__device__ __noinline__ void batch_fold_chain(const BatchFilter &F, FilterCtl *ctl,
                                              const BatchPool &P, BatchLds<Inv> &S)
{
    const uint64_t T = (F.n + kFT - 1) / kFT;
    int bad;
    const Seg tail = S.tail;
    const double tot = fold_chain<kBlock>(P.tpre + F.off / kFT, nullptr, T, tail, P.spec + F.off,
                                          P.gval + F.off, S.r, S.v, &bad);
    if (threadIdx.x == 0) {
        ctl->fold_total = tot;
        ctl->specials = tail.cnt;
        if (bad) atomicOr(&ctl->err, (int)kErrFoldClass);
    }
}

// exact values, inversion, replication; the zero/surplus rank scan
template <typename Inv>
__device__ __noinline__ void batch_fold_apply(const BatchFilter &F, FilterCtl *ctl,
                                              const BatchPool &P, BatchLds<Inv> &S)
{
    const int tid = threadIdx.x;
    const uint64_t N = F.n, T = (N + kFT - 1) / kFT;
    XWeight xs{P.lw + F.off, 0.0, 0.0};
    xs.bind(ctl);
    Inv inv;
    inv.bind(ctl);
    const FoldOut fo{P.anc + F.off, P.zl + F.off, P.sl + F.off, nullptr, nullptr};
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        fold_tile_apply(ctl, xs, inv, tt, N, P.ts[F.off / kFT + tt], P.tpre[F.off / kFT + tt],
                        tt + 1 < T ? P.tpre[F.off / kFT + tt + 1] : seg_identity(), tt + 1 < T,
                        P.gval + F.off, fo,
                        P.rep32 + F.off, P.rep8 + F.off, P.tzs + F.off / kFT + tt, S.L);
    __syncthreads();
    uint64_t *zs = P.tzs + F.off / kFT;
    uint64_t carry = 0;
#pragma unroll 1
    for (uint64_t c0 = 0; c0 < T; c0 += kBlock) {
        const uint64_t i = c0 + tid;
        const uint64_t a = i < T ? zs[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan<kBlock>(a, AddOp(), (uint64_t)0, S.L.u64, &tot);
        if (i < T) zs[i] = carry + ex;  // in place: exclusive zero/surplus ranks
        carry += tot;
    }
    if (tid == 0) {
        S.zeros = carry >> 32;
        ctl->zeros = carry >> 32;
    }
    __syncthreads();
}

// cfrp_trans layout, the zero-slot fill, engine bookkeeping
template <typename Inv>
__device__ __noinline__ void batch_fold_layout(const BatchFilter &F, FilterCtl *ctl,
                                               const BatchPool &P, BatchLds<Inv> &S)
{
    const int tid = threadIdx.x;
    const uint64_t N = F.n, T = (N + kFT - 1) / kFT;
    uint32_t *anc = P.anc + F.off;
    // encoded ancestors (zero slots resolved through the surplus list by the
    // next move), as in the tick: no zero-fill loop
    const FoldOut fo{anc, P.zl + F.off, P.sl + F.off, nullptr, nullptr, nullptr, true};
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        fold_tile_layout(ctl, tt, N, P.rep32 + F.off, P.rep8 + F.off, P.tzs[F.off / kFT + tt],
                         fo, S.L);
    __syncthreads();
    if (tid == 0)
        resample_epilogue(ctl, P.hist + F.hist_off, SchemeOf<Inv>::value == kSystematic, N);
}

// One step t of filter f.
template <typename Real, typename Inv>
__device__ void batch_step(const BatchFilter &F, uint32_t f, const BatchPool &P,
                           const CvParams &p, const LgParams &lp, uint64_t t, BatchLds<Inv> &S)
{
    FilterCtl *ctl = P.ctl + f;
    if (F.model == kModelLg)
        batch_move_lg<Real, Inv>(F, f, P, lp, t, S);
    else
        batch_move_cv<Real, Inv>(F, f, P, p, t, S);
    if (S.flag) {
        batch_fold_sums(F, ctl, P, S);
        batch_fold_segs(F, ctl, P, S);
        batch_fold_chain(F, ctl, P, S);
        batch_fold_apply(F, ctl, P, S);
        batch_fold_layout(F, ctl, P, S);
    }
    if (t + 1 == F.horizon) {
        __syncthreads();
        batch_final_monitor<Real, Inv>(F, f, ctl, P, lp, P.hist + F.hist_off + t, S);
    }
    if (threadIdx.x == 0) {
        P.next_t[f] = t + 1;
        if (ctl->err) atomicOr(P.err_any, ctl->err);
    }
    __syncthreads();
}

// Persistent: workgroups take filters from the queue until it is empty; each
// runs up to max_steps steps of its filter (as many as it has observations).
template <typename Real, typename Inv>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_batch_run(const BatchFilter *filters,
                                                      const uint32_t *order, uint32_t nf,
                                                      BatchPool P, CvParams p, LgParams lp,
                                                      uint64_t max_steps, unsigned *queue)
{
    __shared__ BatchLds<Inv> S;
    for (;;) {
        if (threadIdx.x == 0) S.job = atomicAdd(queue, 1u);
        __syncthreads();
        const unsigned j = S.job;
        __syncthreads();
        if (j >= nf) return;
        const uint32_t f = order[j];
        const BatchFilter F = filters[f];
        const uint64_t t0 = P.next_t[f];
        uint64_t t1 = P.avail[f] < F.horizon ? P.avail[f] : F.horizon;
        if (t1 > t0 + max_steps) t1 = t0 + max_steps;
#pragma unroll 1
        for (uint64_t t = t0; t < t1; ++t) batch_step<Real, Inv>(F, f, P, p, lp, t, S);
    }
}

__global__ void k_batch_queue_reset(unsigned *queue)
{
    if (threadIdx.x == 0) *queue = 0;
}

void launch_batch_run(bool fp64, int scheme, const BatchFilter *filters, const uint32_t *order,
                      uint32_t nf, const BatchPool &P, const CvParams &p, const LgParams &lg,
                      uint64_t max_steps, unsigned *queue, unsigned grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_queue_reset, dim3(1), dim3(64), 0, st, queue);
    const dim3 g(grid ? grid : 1), b(kBlock);
    if (fp64) {
        if (scheme == kStratified)
            hipLaunchKernelGGL((k_batch_run<double, InvStratified>), g, b, 0, st, filters, order,
                               nf, P, p, lg, max_steps, queue);
        else
            hipLaunchKernelGGL((k_batch_run<double, InvSystematic>), g, b, 0, st, filters, order,
                               nf, P, p, lg, max_steps, queue);
    } else {
        if (scheme == kStratified)
            hipLaunchKernelGGL((k_batch_run<float, InvStratified>), g, b, 0, st, filters, order,
                               nf, P, p, lg, max_steps, queue);
        else
            hipLaunchKernelGGL((k_batch_run<float, InvSystematic>), g, b, 0, st, filters, order,
                               nf, P, p, lg, max_steps, queue);
    }
}

}  // namespace vsmcx
