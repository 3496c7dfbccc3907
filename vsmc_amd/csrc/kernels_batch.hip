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
#include "cv_model.hpp"
#include "fold_tiles.hpp"
#include "launch.hpp"

namespace vsmcx {

template <typename Inv>
struct BatchLds {
    FoldLdsT<Inv::kWindow> L;
    double ts[kBatchMaxTiles];
    Seg agg[kBatchMaxTiles];
    Seg pre[kBatchMaxTiles];
    uint64_t zs[kBatchMaxTiles];
    uint64_t zsx[kBatchMaxTiles];
    uint32_t cnt[kBatchMaxTiles];
    SpecRec r[kBlock];
    double v[kBlock];
    Part5 p5[kBlock / kWave];
    Seg tail;
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
// last step (core/monitor.hpp:274-302).
template <typename Real, typename Inv>
__device__ void batch_final_monitor(const FilterCtl *ctl, const typename StateOf<Real>::type *st,
                                    const uint32_t *anc, const double *lw, uint64_t N,
                                    StepRow *row, BatchLds<Inv> &S)
{
    const int prev = ctl->prev_resampled;
    const double m = ctl->m, c = ctl->coeff, invN = 1.0 / (double)N;
    Part5 v{0, 0, 0, 0};
#pragma unroll 1
    for (uint64_t i = threadIdx.x; i < N; i += kBlock) {
        const uint64_t src = prev ? (uint64_t)anc[i] : i;
        const auto s = st[src < N ? src : i];
        const double w = prev ? invN : exp(lw[i] - m) * c;
        v.mx += w * (double)s.x;
        v.my += w * (double)s.y;
    }
    v = block_reduce<kBlock>(v, Part5Add(), S.p5);
    if (threadIdx.x == 0) {
        row->mean_x = v.mx;
        row->mean_y = v.my;
    }
}

// The phases of a step are separate (not inlined) functions: each gets its own
// register allocation, and what crosses phases lives in LDS, so the kernel
// keeps the occupancy of its largest phase instead of the sum of all.

// move, likelihood, log-sum-exp, decision (thread 0 writes the scalars)
template <typename Real, typename Inv>
__device__ __noinline__ void batch_move(const BatchFilter &F, uint32_t f, const BatchPool &P,
                                        const CvParams &p, uint64_t t, BatchLds<Inv> &S)
{
    typedef typename StateOf<Real>::type S4;
    const int tid = threadIdx.x;
    FilterCtl *ctl = P.ctl + f;
    StepRow *hist = P.hist + F.hist_off;
    const uint64_t N = F.n;
    S4 *sa = (S4 *)P.st0 + F.off, *sb = (S4 *)P.st1 + F.off;
    const uint32_t cur = P.cur[f];
    const S4 *in = cur ? sb : sa;
    S4 *out = cur ? sa : sb;
    double *lw = P.lw + F.off;
    const uint32_t *anc = P.anc + F.off;
    CvModel<Real> md;
    md.load(p);
    const int prev = t > 0 ? ctl->prev_resampled : 0;
    const double m_old = ctl->m, c_old = ctl->coeff;
    const double invN = 1.0 / (double)N;
    const Real ox = (Real)P.obs[2 * (F.obs_off + t)], oy = (Real)P.obs[2 * (F.obs_off + t) + 1];
    const uint32_t k0 = (uint32_t)F.seed, k1 = (uint32_t)(F.seed >> 32);
    double m = -INFINITY, se = 0, se2 = 0, mx = 0, my = 0;
#pragma unroll 1
    for (uint64_t i = tid; i < N; i += kBlock) {
        S4 s;
        double lwo = 0;
        if (t > 0) {
            uint64_t src = prev ? (uint64_t)anc[i] : i;
            src = src < N ? src : i;
            s = in[src];
            if (!prev) lwo = lw[i];
            const double w = prev ? invN : exp(lwo - m_old) * c_old;
            mx += w * (double)s.x;
            my += w * (double)s.y;
        }
        Real z[4];
        cv_normals(z, k0, k1, i, t);
        cv_advance(md, s, z, t);
        const double lwn = lwo + (double)md.llh(s.x, s.y, ox, oy);
        out[i] = s;
        lw[i] = lwn;
        lse_add(m, se, se2, lwn);
    }
    const double M = block_reduce<kBlock>(m, MaxOp(), S.L.red);
    const double sc = (m == -INFINITY) ? 0.0 : exp(m - M);
    Part5 v{se * sc, se2 * sc * sc, mx, my};
    v = block_reduce<kBlock>(v, Part5Add(), S.p5);
    if (tid == 0) {
        cv_step_scalars(ctl, M, v.se, v.se2, v.mx, v.my, hist, t);
        P.cur[f] = 1 - cur;
        S.flag = ctl->resample;
        if (ctl->resample) resample_prologue(ctl, SchemeOf<Inv>::value);
    }
    __syncthreads();
}

// tile sums and their exclusive scan (approximate tile starts)
template <typename Inv>
__device__ __noinline__ void batch_fold_sums(const BatchFilter &F, FilterCtl *ctl,
                                             const BatchPool &P, BatchLds<Inv> &S)
{
    const int tid = threadIdx.x;
    const uint64_t N = F.n, T = (N + kFT - 1) / kFT;
    XWeight xs{P.lw + F.off, 0.0, 0.0};
    xs.bind(ctl);
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt) {
        const double s = fold_tile_sum(xs, tt * (uint64_t)kFT, N, S.L.red);
        if (tid == 0) S.ts[tt] = s;
    }
    __syncthreads();
    const double a = tid < (int)T ? S.ts[tid] : 0.0;
    const double ex = block_excl_scan<kBlock>(a, AddOp(), 0.0, S.L.red, (double *)nullptr);
    if (tid < (int)T) S.ts[tid] = ex;
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
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        fold_tile_seg(xs, tt * (uint64_t)kFT, N, S.ts[tt], P.spec + F.off + tt * (uint64_t)kFT,
                      &S.agg[tt], S.L);
    __syncthreads();
    const Seg a = tid < (int)T ? S.agg[tid] : seg_identity();
    const Seg ex = block_excl_scan<kBlock>(a, SegOp(), seg_identity(), S.L.seg, (Seg *)nullptr);
    if (tid < (int)T) {
        S.pre[tid] = ex;
        S.cnt[tid] = (uint32_t)ex.cnt;
    }
    if (tid == kBlock - 1) S.tail = seg_op(ex, a);
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
    const double tot = fold_chain<kBlock>(S.pre, S.cnt, T, tail, P.spec + F.off, P.gval + F.off,
                                          S.r, S.v, &bad);
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
        fold_tile_apply(ctl, xs, inv, tt, N, S.ts[tt], S.pre[tt], P.gval + F.off, fo,
                        P.rep32 + F.off, &S.zs[tt], S.L);
    __syncthreads();
    const uint64_t a = tid < (int)T ? S.zs[tid] : 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan<kBlock>(a, AddOp(), (uint64_t)0, S.L.u64, &total);
    if (tid < (int)T) S.zsx[tid] = ex;
    if (tid == 0) {
        S.zeros = total >> 32;
        ctl->zeros = total >> 32;
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
    const FoldOut fo{anc, P.zl + F.off, P.sl + F.off, nullptr, nullptr};
#pragma unroll 1
    for (uint64_t tt = 0; tt < T; ++tt)
        fold_tile_layout(ctl, tt, N, P.rep32 + F.off, S.zsx[tt], fo, S.L);
    __syncthreads();
    const uint64_t z = S.zeros < N ? S.zeros : N;
#pragma unroll 1
    for (uint64_t k = tid; k < z; k += kBlock) {
        const uint32_t to = fo.zl[k], from = fo.sl[k];
        if (to < N && from < N)
            anc[to] = from;
        else
            atomicOr(&ctl->err, (int)kErrBounds);
    }
    __syncthreads();
    if (tid == 0)
        resample_epilogue(ctl, P.hist + F.hist_off, SchemeOf<Inv>::value == kSystematic, N);
}

// One step t of filter f.
template <typename Real, typename Inv>
__device__ void batch_step(const BatchFilter &F, uint32_t f, const BatchPool &P,
                           const CvParams &p, uint64_t t, BatchLds<Inv> &S)
{
    FilterCtl *ctl = P.ctl + f;
    batch_move<Real, Inv>(F, f, P, p, t, S);
    if (S.flag) {
        batch_fold_sums(F, ctl, P, S);
        batch_fold_segs(F, ctl, P, S);
        batch_fold_chain(F, ctl, P, S);
        batch_fold_apply(F, ctl, P, S);
        batch_fold_layout(F, ctl, P, S);
    }
    if (t + 1 == F.horizon) {
        __syncthreads();
        typedef typename StateOf<Real>::type S4;
        const S4 *cur_state = (P.cur[f] ? (const S4 *)P.st1 : (const S4 *)P.st0) + F.off;
        batch_final_monitor<Real, Inv>(ctl, cur_state, P.anc + F.off, P.lw + F.off, F.n,
                                       P.hist + F.hist_off + t, S);
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
                                                      BatchPool P, CvParams p, uint64_t max_steps,
                                                      unsigned *queue)
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
        for (uint64_t t = t0; t < t1; ++t) batch_step<Real, Inv>(F, f, P, p, t, S);
    }
}

__global__ void k_batch_queue_reset(unsigned *queue)
{
    if (threadIdx.x == 0) *queue = 0;
}

void launch_batch_run(bool fp64, int scheme, const BatchFilter *filters, const uint32_t *order,
                      uint32_t nf, const BatchPool &P, const CvParams &p, uint64_t max_steps,
                      unsigned *queue, unsigned grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_queue_reset, dim3(1), dim3(64), 0, st, queue);
    const dim3 g(grid ? grid : 1), b(kBlock);
    if (fp64) {
        if (scheme == kStratified)
            hipLaunchKernelGGL((k_batch_run<double, InvStratified>), g, b, 0, st, filters, order,
                               nf, P, p, max_steps, queue);
        else
            hipLaunchKernelGGL((k_batch_run<double, InvSystematic>), g, b, 0, st, filters, order,
                               nf, P, p, max_steps, queue);
    } else {
        if (scheme == kStratified)
            hipLaunchKernelGGL((k_batch_run<float, InvStratified>), g, b, 0, st, filters, order,
                               nf, P, p, max_steps, queue);
        else
            hipLaunchKernelGGL((k_batch_run<float, InvSystematic>), g, b, 0, st, filters, order,
                               nf, P, p, max_steps, queue);
    }
}

}  // namespace vsmcx
