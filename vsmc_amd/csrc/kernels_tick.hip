// kernels_tick.hip -- one serving tick: one step of every job in the tick,
// each phase one grid launch over the tiles of all of them.
//
// A tick is the single-filter pipeline (kernels_cv.hip, kernels_fold.hip)
// made segmented: the tile work of every job is one flat list of
// (job, tile) items, and the per-job reductions and scans run one workgroup
// per job.  The per-tile and per-job bodies are the single-filter ones
// (cv_tiles.hpp, fold_tiles.hpp) on the same tiles in the same order, so a
// served CV job computes exactly what vsmcx_filter computes, and no job's
// arithmetic depends on the other jobs in the tick.  Jobs that do not
// resample this step leave the fold launches at once.
#include "cv_tiles.hpp"
#include "fold_tiles.hpp"
#include "launch.hpp"
#include "lg_tiles.hpp"

namespace vsmcx {

// Per-job kernels: the scans of a job's tile summaries (a job holds at most
// kBatchMaxTiles = 512 tiles) run in workgroups of kJobBlock threads; the
// reductions of tile partials use one wave.
constexpr int kJobBlock = 256;

__device__ __forceinline__ uint64_t tiles_of_job(const BatchFilter &F)
{
    return (F.n + kFT - 1) / kFT;
}

// ---- move (all jobs' tiles) ---------------------------------------------------------
// kCvOnly: every job of the launch runs the CV model (the batch), so the LG
// body is not compiled in.  FP32 with both bodies: compiled for five waves per
// SIMD.
template <typename Real, bool kCvOnly>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(std::is_same<Real, float>::value && !kCvOnly ? 5 : 1)))
void k_tick_move(const TickItems items,
                                                      const BatchFilter *filters, BatchPool P,
                                                      CvParams cv, LgParams lg)
{
    const TickItem it = items[blockIdx.x];
    const BatchFilter F = filters[it.slot];
    FilterCtl *ctl = P.ctl + it.slot;
    const uint64_t t = P.next_t[it.slot];
    const uint32_t cur = P.cur[it.slot];
    const uint32_t k0 = (uint32_t)F.seed, k1 = (uint32_t)(F.seed >> 32);
    double *part = P.tpart + 5 * (F.off / kFT + it.tile);
    if (!kCvOnly && F.model == kModelLg) {
        Real *sa = (Real *)P.st0 + F.soff, *sb = (Real *)P.st1 + F.soff;
        lg_move_tile<Real>(ctl, cur ? sb : sa, cur ? sa : sb, P.anc + F.off, P.sl + F.off, P.lw + F.off,
                           P.obs + F.obs_off + (uint64_t)lg.dim * t, lg, k0, k1, t, F.n, it.tile,
                           part);
    } else {
        typedef typename StateOf<Real>::type S4;
        S4 *sa = (S4 *)((Real *)P.st0 + F.soff), *sb = (S4 *)((Real *)P.st1 + F.soff);
        cv_move_tile<Real>(ctl, cur ? sb : sa, cur ? sa : sb, P.anc + F.off, P.sl + F.off, P.lw + F.off,
                           P.obs + F.obs_off, cv, k0, k1, t, F.n, it.tile, part);
    }
}

// ---- per job: reduction, decision, resample prologue -----------------------------
__global__ __launch_bounds__(kWave) void k_tick_reduce(const uint32_t *jobs,
                                                      const BatchFilter *filters, BatchPool P,
                                                      int scheme)
{
    const uint32_t slot = jobs[blockIdx.x];
    const BatchFilter F = filters[slot];
    FilterCtl *ctl = P.ctl + slot;
    const uint64_t t = P.next_t[slot];
    cv_reduce_tiles(ctl, P.tpart + 5 * (F.off / kFT), tiles_of_job(F), P.hist + F.hist_off, t);
    if (threadIdx.x == 0) {
        P.cur[slot] ^= 1u;
        if (ctl->resample) resample_prologue(ctl, scheme);
    }
}

// ---- fold passes -------------------------------------------------------------------
// The tile starts come from the move's tile partials (TileSumsFromLse), as in
// the single filter: no tile-sum pass.
__global__ __launch_bounds__(kJobBlock) void k_tick_scan_sum(const uint32_t *jobs,
                                                             const BatchFilter *filters,
                                                             BatchPool P)
{
    const uint32_t slot = jobs[blockIdx.x];
    const FilterCtl *ctl = P.ctl + slot;
    if (!ctl->resample) return;
    __shared__ double s_red[kJobBlock / kWave];
    const BatchFilter F = filters[slot];
    const uint64_t o = F.off / kFT;
    const TileSumsFromLse in{P.tpart + 5 * o, ctl->m, ctl->coeff};
    wg_scan_b<kJobBlock>(in, P.ts + o, tiles_of_job(F), AddOp(), 0.0, s_red);
}

// kCache: x left in P.wx for the apply pass (the stratified inversion); the
// systematic apply computes x from the log-weights again.
template <bool kCache>
__global__ __launch_bounds__(kBlock) void k_tick_seg(const TickItems items,
                                                     const BatchFilter *filters, BatchPool P)
{
    const TickItem it = items[blockIdx.x];
    FilterCtl *ctl = P.ctl + it.slot;
    if (!ctl->resample) return;
    __shared__ FoldLdsT<false> L;
    const BatchFilter F = filters[it.slot];
    XWeight xs{P.lw + F.off, 0.0, 0.0};
    xs.bind(ctl);
    const uint64_t o = F.off / kFT + it.tile;
    fold_tile_seg(xs, it.tile * (uint64_t)kFT, F.n, P.ts[o], P.spec + F.off + it.tile * (uint64_t)kFT,
                  P.tagg + o, L, P.thr + o * kBlock, P.thr_fb + o * kBlock,
                  kCache ? P.wx + F.off : nullptr);
}

__global__ __launch_bounds__(kJobBlock) void k_tick_scan_seg(const uint32_t *jobs,
                                                             const BatchFilter *filters,
                                                             BatchPool P)
{
    const uint32_t slot = jobs[blockIdx.x];
    FilterCtl *ctl = P.ctl + slot;
    if (!ctl->resample) return;
    __shared__ Seg s_seg[kJobBlock / kWave];
    __shared__ uint32_t s_cnt[kBatchMaxTiles + 1];
    __shared__ SpecRec s_r[kJobBlock];
    __shared__ double s_v[kJobBlock];
    const BatchFilter F = filters[slot];
    const uint64_t o = F.off / kFT, T = tiles_of_job(F);
    const Seg tail = wg_scan_b<kJobBlock>(P.tagg + o, P.tpre + o, T, SegOp(), seg_identity(), s_seg,
                                          [&](uint64_t t, const Seg &p) { s_cnt[t] = (uint32_t)p.cnt; });
    __syncthreads();
    int bad;
    const double tot = fold_chain<kJobBlock>(P.tpre + o, s_cnt, T, tail, P.spec + F.off,
                                             P.gval + F.off, s_r, s_v, &bad);
    if (threadIdx.x == 0) {
        ctl->fold_total = tot;
        ctl->specials = tail.cnt;
        if (bad) atomicOr(&ctl->err, (int)kErrFoldClass);
    }
}

template <typename Inv>
__global__ __launch_bounds__(kBlock) void k_tick_apply(const TickItems items,
                                                       const BatchFilter *filters, BatchPool P)
{
    const TickItem it = items[blockIdx.x];
    FilterCtl *ctl = P.ctl + it.slot;
    if (!ctl->resample) return;
    __shared__ FoldLdsT<Inv::kWindowTmp> L;
    const BatchFilter F = filters[it.slot];
    // the weights (cached by the segment pass for the stratified inversion)
    // and the segment pass's per-thread prefixes
    typedef typename std::conditional<std::is_same<Inv, InvSystematic>::value, XWeight,
                                      XCached<XWeight>>::type XS;
    XS xs;
    if constexpr (std::is_same<Inv, InvSystematic>::value)
        xs = XWeight{P.lw + F.off, 0.0, 0.0};
    else
        xs = XCached<XWeight>{P.wx + F.off, XWeight{P.lw + F.off, 0.0, 0.0}};
    xs.bind(ctl);
    Inv inv;
    inv.bind(ctl);
    const uint64_t o = F.off / kFT + it.tile;
    const FoldOut fo{P.anc + F.off, P.zl + F.off, P.sl + F.off, nullptr, nullptr};
    const bool has_next = it.tile + 1 < tiles_of_job(F);
    fold_tile_apply(ctl, xs, inv, it.tile, F.n, P.ts[o], P.tpre[o],
                    has_next ? P.tpre[o + 1] : seg_identity(), has_next, P.gval + F.off, fo,
                    P.rep32 + F.off, P.rep8 + F.off,
                    P.tzs + o, L, P.thr + o * kBlock, P.thr_fb + o * kBlock);
}

__global__ __launch_bounds__(kJobBlock) void k_tick_scan_cnt(const uint32_t *jobs,
                                                             const BatchFilter *filters,
                                                             BatchPool P)
{
    const uint32_t slot = jobs[blockIdx.x];
    FilterCtl *ctl = P.ctl + slot;
    if (!ctl->resample) return;
    __shared__ uint64_t s_red[kJobBlock / kWave];
    const BatchFilter F = filters[slot];
    const uint64_t o = F.off / kFT;
    const uint64_t tot = wg_scan_b<kJobBlock>(P.tzs + o, P.tzsx + o, tiles_of_job(F), AddOp(),
                                              (uint64_t)0, s_red);
    if (threadIdx.x == 0) ctl->zeros = tot >> 32;
}

__global__ __launch_bounds__(kBlock) void k_tick_layout(const TickItems items,
                                                        const BatchFilter *filters, BatchPool P)
{
    const TickItem it = items[blockIdx.x];
    FilterCtl *ctl = P.ctl + it.slot;
    if (!ctl->resample) return;
    __shared__ FoldLdsT<false> L;
    const BatchFilter F = filters[it.slot];
    // encoded ancestors: the move of the next step reads zero slots' sources
    // from the surplus list, so there is no zero-fill pass
    const FoldOut fo{P.anc + F.off, P.zl + F.off, P.sl + F.off, nullptr, nullptr, nullptr, true};
    fold_tile_layout(ctl, it.tile, F.n, P.rep32 + F.off, P.rep8 + F.off,
                     P.tzsx[F.off / kFT + it.tile], fo, L);
}

__global__ void k_tick_epilogue(const uint32_t *jobs, uint32_t njobs, const BatchFilter *filters,
                                BatchPool P, int scheme)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const uint32_t slot = jobs[j];
    FilterCtl *ctl = P.ctl + slot;
    const BatchFilter F = filters[slot];
    if (ctl->resample) resample_epilogue(ctl, P.hist + F.hist_off, scheme == kSystematic, F.n);
}

// ---- the last step's monitor, for jobs that finish in this tick ----------------------
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_tick_monitor(const TickItems items,
                                                         const BatchFilter *filters, BatchPool P,
                                                         LgParams lg)
{
    const TickItem it = items[blockIdx.x];
    const BatchFilter F = filters[it.slot];
    const FilterCtl *ctl = P.ctl + it.slot;
    if (P.next_t[it.slot] + 1 != F.horizon) return;
    const uint32_t cur = P.cur[it.slot];  // flipped by the reduce: the new state
    double *part = P.tpart + 5 * (F.off / kFT + it.tile);
    if (F.model == kModelLg) {
        const Real *st = (cur ? (const Real *)P.st1 : (const Real *)P.st0) + F.soff;
        lg_monitor_tile<Real>(ctl, st, lg.dim, P.anc + F.off, P.sl + F.off, P.lw + F.off, F.n, it.tile, part);
    } else {
        typedef typename StateOf<Real>::type S4;
        const S4 *st = (const S4 *)((cur ? (const Real *)P.st1 : (const Real *)P.st0) + F.soff);
        cv_monitor_tile<Real>(ctl, st, P.anc + F.off, P.sl + F.off, P.lw + F.off, F.n, it.tile, part);
    }
}

// One wave per job: the final monitor's reduction for jobs at their last
// step, then the step counter and the error summary.
__global__ __launch_bounds__(kWave) void k_tick_finish(const uint32_t *jobs,
                                                       const BatchFilter *filters, BatchPool P)
{
    const uint32_t slot = jobs[blockIdx.x];
    const BatchFilter F = filters[slot];
    const uint64_t t = P.next_t[slot];
    if (t + 1 == F.horizon)
        cv_monitor_reduce(P.ctl + slot, P.tpart + 5 * (F.off / kFT), tiles_of_job(F),
                          P.hist + F.hist_off);
    if (threadIdx.x == 0) {
        P.next_t[slot] = t + 1;
        const int e = P.ctl[slot].err;
        if (e) atomicOr(P.err_any, e);
    }
}

// One workgroup per slot initialisation, per observation run, and per 64
// availability counts; every record is written by exactly one workgroup.
__global__ void __launch_bounds__(64) k_tick_upload(TickUpload u, BatchFilter *filters, BatchPool P)
{
    const uint32_t b = blockIdx.x;
    if (b < u.n_init) {
        const SlotInit &r = u.init[b];
        static_assert(sizeof(BatchFilter) % 4 == 0 && sizeof(FilterCtl) % 4 == 0, "word copies");
        const uint32_t *fs = reinterpret_cast<const uint32_t *>(&r.f);
        const uint32_t *cs = reinterpret_cast<const uint32_t *>(&r.c);
        uint32_t *fd = reinterpret_cast<uint32_t *>(filters + r.slot);
        uint32_t *cd = reinterpret_cast<uint32_t *>(P.ctl + r.slot);
        for (uint32_t w = threadIdx.x; w < sizeof(BatchFilter) / 4; w += 64) fd[w] = fs[w];
        for (uint32_t w = threadIdx.x; w < sizeof(FilterCtl) / 4; w += 64) cd[w] = cs[w];
        if (threadIdx.x == 0) {
            P.next_t[r.slot] = 0;
            P.cur[r.slot] = 0;
        }
        return;
    }
    if (b < u.n_init + u.n_runs) {
        const ObsRun &r = u.runs[b - u.n_init];
        double *obs = const_cast<double *>(P.obs);  // read-only to the step kernels
        for (uint32_t k = threadIdx.x; k < r.count; k += 64) obs[r.dst + k] = u.pool[r.src + k];
        return;
    }
    const uint32_t k = (b - u.n_init - u.n_runs) * 64 + threadIdx.x;
    if (k < u.n_avail) const_cast<uint64_t *>(P.avail)[u.avail[k].slot] = u.avail[k].avail;
}

void launch_tick_upload(const TickUpload &u, BatchFilter *filters, const BatchPool &P,
                        hipStream_t st)
{
    const uint32_t g = u.n_init + u.n_runs + (u.n_avail + 63) / 64;
    if (!g) return;
    hipLaunchKernelGGL(k_tick_upload, dim3(g), dim3(64), 0, st, u, filters, P);
}

void launch_tick(bool fp64, int scheme, const TickItems &items, uint32_t nitems,
                 const uint32_t *jobs, uint32_t njobs, const BatchFilter *filters,
                 const BatchPool &P, const CvParams &cv, const LgParams &lg, hipStream_t st,
                 bool cv_only)
{
    if (!nitems || !njobs) return;
    const dim3 gi(nitems), gj(njobs), b(kBlock), bs(kJobBlock), bw(kWave);
    const dim3 gsmall((njobs + 255) / 256), bsmall(256);
    if (fp64 && cv_only)
        hipLaunchKernelGGL((k_tick_move<double, true>), gi, b, 0, st, items, filters, P, cv, lg);
    else if (fp64)
        hipLaunchKernelGGL((k_tick_move<double, false>), gi, b, 0, st, items, filters, P, cv, lg);
    else if (cv_only)
        hipLaunchKernelGGL((k_tick_move<float, true>), gi, b, 0, st, items, filters, P, cv, lg);
    else
        hipLaunchKernelGGL((k_tick_move<float, false>), gi, b, 0, st, items, filters, P, cv, lg);
    hipLaunchKernelGGL(k_tick_reduce, gj, bw, 0, st, jobs, filters, P, scheme);
    hipLaunchKernelGGL(k_tick_scan_sum, gj, bs, 0, st, jobs, filters, P);
    if (scheme == kStratified)
        hipLaunchKernelGGL(k_tick_seg<true>, gi, b, 0, st, items, filters, P);
    else
        hipLaunchKernelGGL(k_tick_seg<false>, gi, b, 0, st, items, filters, P);
    hipLaunchKernelGGL(k_tick_scan_seg, gj, bs, 0, st, jobs, filters, P);
    if (scheme == kStratified)
        hipLaunchKernelGGL(k_tick_apply<InvStratified>, gi, b, 0, st, items, filters, P);
    else
        hipLaunchKernelGGL(k_tick_apply<InvSystematic>, gi, b, 0, st, items, filters, P);
    hipLaunchKernelGGL(k_tick_scan_cnt, gj, bs, 0, st, jobs, filters, P);
    hipLaunchKernelGGL(k_tick_layout, gi, b, 0, st, items, filters, P);
    hipLaunchKernelGGL(k_tick_epilogue, gsmall, bsmall, 0, st, jobs, njobs, filters, P, scheme);
    if (fp64)
        hipLaunchKernelGGL(k_tick_monitor<double>, gi, b, 0, st, items, filters, P, lg);
    else
        hipLaunchKernelGGL(k_tick_monitor<float>, gi, b, 0, st, items, filters, P, lg);
    hipLaunchKernelGGL(k_tick_finish, gj, bw, 0, st, jobs, filters, P);
}

}  // namespace vsmcx
