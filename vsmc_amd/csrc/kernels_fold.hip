// kernels_fold.hip -- exact-fold resampling on the device.
//
// The passes below replace vSMC's sequential
//   inversion  (include/vsmc/resample/common.hpp:57-84)  and
//   cfrp_trans (include/vsmc/resample/common.hpp:87-118),
// producing the replication counts, the keep-in-place ancestor layout and the
// lists the zero-slot fill (k_zero_fill) consumes.  The running weight sum is
// the exact left fold vSMC computes (see fold_core.hpp), so given the same
// weights and uniforms the ancestors are the reference's, at any N.
#include "fold_tiles.hpp"
#include "launch.hpp"

namespace vsmcx {

// ------------------------------------------------------------------------
// the grid passes over the tiles of one particle system
// ------------------------------------------------------------------------
// ---- pass 1: tile sums of x ---------------------------------------------------
template <typename XSrc>
__global__ __launch_bounds__(kBlock) void k_fold_tsum(FilterCtl *ctl, XSrc xs, uint64_t M,
                                                      FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kBlock / kWave];
    xs.bind(ctl);
    const double s = fold_tile_sum(xs, blockIdx.x * (uint64_t)kFT, M, s_red, fs.wx);
    if (threadIdx.x == 0) fs.tsum[blockIdx.x] = s;
}

// ---- pass 2: approximate tile starts (one workgroup) ---------------------------
__global__ __launch_bounds__(kScanBlock) void k_fold_scan_sum(FilterCtl *ctl, FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kScanBlock / kWave];
    wg_scan<double>(fs.tsum, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
}

// ---- pass 2 from the move's partials: tile starts of the filter's weights -------
// The filter's move leaves each tile's {max, sum exp(lw - max)}; the tile sums
// of x_i = exp(lw_i - m) coeff follow from them (TileSumsFromLse), so the
// tile-sum pass is not needed.
__global__ __launch_bounds__(kScanBlock) void k_fold_scan_sum_lse(FilterCtl *ctl,
                                                                  const double *lse_part,
                                                                  FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kScanBlock / kWave];
    const TileSumsFromLse in{lse_part, ctl->m, ctl->coeff};
    wg_scan<double>(in, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
}

// ---- pass 3: classification, tile segment summaries, special records -----------
// wx_out: where the pass leaves the x values for the apply pass (when no
// tile-sum pass wrote them).
template <typename XSrc>
__global__ __launch_bounds__(kBlock) void k_fold_seg(FilterCtl *ctl, XSrc xs, uint64_t M,
                                                     FoldScratch fs, double *wx_out)
{
    if (!ctl->resample) return;
    __shared__ FoldLdsT<false> L;
    xs.bind(ctl);
    const uint64_t t = blockIdx.x;
    fold_tile_seg(xs, t * (uint64_t)kFT, M, fs.ts[t], fs.spec + t * (uint64_t)kFT, fs.agg + t, L,
                  fs.thr ? fs.thr + t * kBlock : nullptr, fs.thr ? fs.thr_fb + t * kBlock : nullptr,
                  wx_out);
}

// ---- pass 4a: segment prefixes within chunks (one wave per chunk) ----------------
// Each chunk of kFoldChunk consecutive tiles: every tile's exclusive segment
// summary within the chunk (pre, and its special count in tcnt), and the
// chunk's summary (cagg).  Spread over many CUs, the loads overlap; a single
// workgroup over every tile summary was bound by the cache-line misses one CU
// keeps in flight.
__global__ __launch_bounds__(kWave) void k_fold_scan_chunks(FilterCtl *ctl, FoldScratch fs)
{
    if (!ctl->resample) return;
    const uint64_t T = fs.tiles;
    const uint64_t c = blockIdx.x;
    const uint64_t t = c * kFoldChunk + threadIdx.x;
    const Seg v = t < T ? fs.agg[t] : seg_identity();
    const Seg incl = wave_incl_scan(v, SegOp());
    Seg ex = lane_up(incl, 1);
    if (threadIdx.x == 0) ex = seg_identity();
    if (t < T) {
        fs.pre[t] = ex;
        fs.tcnt[t] = (uint32_t)ex.cnt;
    }
    if (threadIdx.x == kWave - 1) fs.cagg[c] = incl;
}

// ---- pass 4b: chunk prefixes and the exact values at the specials -----------------
// One workgroup: the exclusive summary of every chunk (the special counts also
// kept in LDS for up to kLdsChunks chunks), then the chain.
constexpr int kLdsChunks = 16384;
static_assert(kFoldChunk == kWave, "one wave scans one chunk");

// The chain over the specials in fold order, with the tile prefixes in two
// levels (chunk prefix cpre, then pre within the chunk): as fold_chain, the
// record of special g is found from the counts (a binary search over the
// chunks, then the chunk's kFoldChunk tile counts, loaded at once).
template <int BLOCK>
__device__ double fold_chain_chunked(const FoldScratch &fs, const uint32_t *ccnt_lds,
                                     uint64_t C, uint64_t T, const Seg &tail, SpecRec *s_r,
                                     double *s_v, int *bad)
{
    const int tid = threadIdx.x;
    const uint64_t S = tail.cnt;
    double a = 0.0;  // running exact value (thread 0)
    *bad = 0;
    for (uint64_t c0 = 0; c0 < S; c0 += BLOCK) {
        const uint64_t g = c0 + tid;
        if (g < S) {
            // last chunk whose count of earlier specials is <= g
            uint64_t lo = 0, hi = C - 1;
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) / 2;
                const uint64_t cc = ccnt_lds ? ccnt_lds[mid] : fs.cpre[mid].cnt;
                if (cc <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const Seg cp = fs.cpre[lo];
            const uint64_t r = g - cp.cnt;  // rank within the chunk
            const uint64_t t0 = lo * kFoldChunk;
            const int len = (int)(T - t0 < (uint64_t)kFoldChunk ? T - t0 : (uint64_t)kFoldChunk);
            // last tile of the chunk with count <= r (counts are non-decreasing, the first 0)
            int j = 0;
#pragma unroll 8
            for (int q = 1; q < len; ++q) j += fs.tcnt[t0 + q] <= r ? 1 : 0;
            const uint64_t t = t0 + j;
            const Seg tp = fs.pre[t];
            const Seg p = seg_op(cp, tp);
            const uint64_t k = r - tp.cnt;
            SpecRec rec = fs.spec[t * (uint64_t)kFT + k];
            if (k == 0) rec.run = tr_compose(p.tr, rec.run);
            s_r[tid] = rec;
        }
        __syncthreads();
        if (tid == 0)
            chain_run(s_r, (int)(S - c0 < (uint64_t)BLOCK ? S - c0 : (uint64_t)BLOCK), a, s_v, bad);
        __syncthreads();
        if (g < S) fs.gval[g] = s_v[tid];
        __syncthreads();
    }
    double tot = a;
    if (tid == 0 && !tr_apply(tail.tr, a, &tot)) {
        *bad = 1;
        tot = a;
    }
    return tot;
}

__global__ __launch_bounds__(kScanBlock) void k_fold_scan_seg(FilterCtl *ctl, FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ Seg s_seg[kScanBlock / kWave];
    __shared__ uint32_t s_cnt[kLdsChunks + 1];  // specials before each chunk
    __shared__ SpecRec s_r[kScanBlock];
    __shared__ double s_v[kScanBlock];
    const uint64_t T = fs.tiles;
    const uint64_t C = (T + kFoldChunk - 1) / kFoldChunk;
    const bool in_lds = C <= (uint64_t)kLdsChunks;
    Seg tail;
    if (C <= (uint64_t)kScanBlock) {
        // one chunk per thread: a single workgroup scan, no serial run per thread
        const uint64_t c = threadIdx.x;
        const Seg v = c < C ? fs.cagg[c] : seg_identity();
        const Seg ex = block_excl_scan<kScanBlock>(v, SegOp(), seg_identity(), s_seg, &tail);
        if (c < C) {
            fs.cpre[c] = ex;
            if (in_lds) s_cnt[c] = (uint32_t)ex.cnt;
        }
    } else {
        tail = wg_scan<Seg>(fs.cagg, fs.cpre, C, SegOp(), seg_identity(), s_seg,
                            [&](uint64_t c, const Seg &p) {
                                if (in_lds) s_cnt[c] = (uint32_t)p.cnt;
                            });
    }
    __syncthreads();  // cpre[] and s_cnt written by other threads
    int bad;
    const double tot =
        fold_chain_chunked<kScanBlock>(fs, in_lds ? s_cnt : nullptr, C, T, tail, s_r, s_v, &bad);
    if (threadIdx.x == 0) {
        ctl->fold_total = tot;
        ctl->specials = tail.cnt;
        if (bad) atomicOr(&ctl->err, (int)kErrFoldClass);
    }
}

// both passes of the segment scan
static void launch_scan_seg(FilterCtl *ctl, FoldScratch &fs, hipStream_t st)
{
    const unsigned C = (unsigned)((fs.tiles + kFoldChunk - 1) / kFoldChunk);
    hipLaunchKernelGGL(k_fold_scan_chunks, dim3(C ? C : 1), dim3(kWave), 0, st, ctl, fs);
    hipLaunchKernelGGL(k_fold_scan_seg, dim3(1), dim3(kScanBlock), 0, st, ctl, fs);
}

// ---- pass 5: exact values, inversion, replication --------------------------------
template <typename XSrc, typename Inv>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(Inv::kMinWaves)))
void k_fold_apply(FilterCtl *ctl, XSrc xs, Inv inv, uint64_t M, FoldScratch fs, FoldOut out)
{
    if (!ctl->resample) return;
    __shared__ FoldLdsT<Inv::kWindowTmp> L;
    xs.bind(ctl);
    inv.bind(ctl);
    const uint64_t t = blockIdx.x;
    const Seg tpre = seg_op(fs.cpre[t / kFoldChunk], fs.pre[t]);
    const bool has_next = t + 1 < (M + kFT - 1) / kFT;
    const Seg next =
        has_next ? seg_op(fs.cpre[(t + 1) / kFoldChunk], fs.pre[t + 1]) : seg_identity();
    fold_tile_apply(ctl, xs, inv, t, M, fs.ts[t], tpre, next, has_next, fs.gval, out, fs.rep32,
                    fs.rep8,
                    fs.zs + t, L, fs.thr ? fs.thr + t * kBlock : nullptr,
                    fs.thr ? fs.thr_fb + t * kBlock : nullptr);
}

// ---- pass 6: zero / surplus ranks of the tiles (one workgroup) ---------------------
// The resample's engine bookkeeping (resample_epilogue) rides along when
// epi.on: n_final is known once the apply pass is done, and nothing after
// this pass reads the engine.
struct Epi {
    StepRow *hist;
    int on;
    int systematic_like;
    const int *uerr = nullptr;  // error bits of the uniforms' fold on the side stream
};

__global__ __launch_bounds__(kScanBlock) void k_fold_scan_cnt(FilterCtl *ctl, FoldScratch fs,
                                                             Epi epi, uint64_t M)
{
    if (!ctl->resample) return;
    __shared__ uint64_t s_red[kScanBlock / kWave];
    const uint64_t tot = wg_scan<uint64_t>(fs.zs, fs.zsx, fs.tiles, AddOp(), (uint64_t)0, s_red);
    if (threadIdx.x == 0) {
        ctl->zeros = tot >> 32;
        if (epi.uerr && *epi.uerr) atomicOr(&ctl->err, *epi.uerr);
        if (epi.on) resample_epilogue(ctl, epi.hist, epi.systematic_like, M);
    }
}

// ---- pass 7: cfrp_trans layout ----------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fold_layout(FilterCtl *ctl, FoldScratch fs,
                                                        FoldOut out, uint64_t M)
{
    if (!ctl->resample) return;
    __shared__ FoldLdsT<false> L;
    const uint64_t t = blockIdx.x;
    fold_tile_layout(ctl, t, M, fs.rep32, fs.rep8, fs.zsx[t], out, L);
}

// ------------------------------------------------------------------------
// zero-slot fill + engine bookkeeping
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_zero_fill(FilterCtl *ctl, const uint32_t *zl,
                                                      const uint32_t *sl, uint32_t *anc,
                                                      uint64_t M)
{
    if (!ctl->resample) return;
    const uint64_t z = ctl->zeros < M ? ctl->zeros : M;
    for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < z;
         k += (uint64_t)gridDim.x * kBlock) {
        const uint32_t to = zl[k], from = sl[k];
        if (to < M && from < M)
            anc[to] = from;
        else
            atomicOr(&ctl->err, (int)kErrBounds);
    }
}

// engine bookkeeping: vSMC's lazy sequences draw min(nn, n_final + 1)
// uniforms (u01.hpp:38-40, common.hpp:76-83); the systematic sequence draws
// one in its constructor (u01.hpp:258-263)
__global__ void k_resample_epilogue(FilterCtl *ctl, StepRow *hist, int systematic_like,
                                    uint64_t M)
{
    if (!ctl->resample || threadIdx.x != 0) return;
    resample_epilogue(ctl, hist, systematic_like, M);
}

// Residual, the asum fold's tile-sum pass: r_i = modf(N W_i, &i_i)
// (residual.hpp:57-61) once per element.  The fractions r_i go to wx and the
// integer parts to rep8 as bytes (255: the part itself in rep32), where the
// scaled fold reads them back (its inversion pass reads every i_i of a tile
// before it writes the tile's replication counts there); the tile sums of r_i
// go to tsum and those of i_i to zs, which the start scan consumes before
// the inversion pass reuses it.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_resid_tsum(FilterCtl *ctl, W w, uint64_t M,
                                                       FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kBlock / kWave];
    __shared__ uint64_t s_u64[kBlock / kWave];
    XResid<W> xr{w, 0.0};
    xr.bind(ctl);
    const uint64_t base = blockIdx.x * (uint64_t)kFT;
    double s = 0;
    uint64_t is = 0;
    // the weights of all the thread's elements first (no branch between the
    // loads), then the splits and stores
    double wv[kFI];
    tile_values(xr.w, base, M, wv);
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const uint64_t i = base + k * kBlock + threadIdx.x;
        double r = 0.0;
        if (i < M) {
            double ip;
            r = modf(xr.nd * wv[k], &ip);
            const uint32_t iu = (uint32_t)ip;
            fs.wx[i] = r;
            fs.rep8[i] = (uint8_t)(iu < 255u ? iu : 255u);
            if (iu >= 255u) fs.rep32[i] = iu;
            is += (uint64_t)ip;
        }
        s += r;
    }
    s = block_reduce<kBlock>(s, AddOp(), s_red);
    is = block_reduce<kBlock>(is, AddOp(), s_u64);
    if (threadIdx.x == 0) {
        fs.tsum[blockIdx.x] = s;
        fs.zs[blockIdx.x] = is;
    }
}

// Residual: tile starts of the fractions (one workgroup), and the total R of
// the integer parts (exact integer sum).
__global__ __launch_bounds__(kScanBlock) void k_fold_scan_sum_resid(FilterCtl *ctl, FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kScanBlock / kWave];
    __shared__ uint64_t s_u64[kScanBlock / kWave];
    wg_scan<double>(fs.tsum, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
    uint64_t r = 0;
    for (uint64_t t = threadIdx.x; t < fs.tiles; t += kScanBlock) r += fs.zs[t];
    r = block_reduce<kScanBlock>(r, AddOp(), s_u64);
    if (threadIdx.x == 0) ctl->r_int = r;
}

// Residual: tile starts of the scaled fractions r_i * rc from the tile sums of
// r_i the asum fold left in tsum (TileSumsScaled).
__global__ __launch_bounds__(kScanBlock) void k_fold_scan_sum_scaled(FilterCtl *ctl,
                                                                     FoldScratch fs)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kScanBlock / kWave];
    const TileSumsScaled in{fs.tsum, ctl->rcoeff};
    wg_scan<double>(in, fs.ts, fs.tiles, AddOp(), 0.0, s_red);
}

// after the residual fold total: rcoeff, nn, delta and (systematic) u0
__global__ void k_residual_prologue(FilterCtl *ctl, int scheme)
{
    if (!ctl->resample || threadIdx.x != 0) return;
    ctl->rcoeff = 1 / ctl->fold_total;
    ctl->nn = ctl->n - ctl->r_int;
    ctl->delta = 1.0 / (double)ctl->nn;
    if (scheme == kResidualSystematic) ctl->u0 = canonical64(tf64_stream_word(ctl->eng, 0)) * ctl->delta;
}

// non-residual prologue: nn = n, delta, u0 (systematic ctor draw)
__global__ void k_plain_prologue(FilterCtl *ctl, int scheme)
{
    if (!ctl->resample || threadIdx.x != 0) return;
    resample_prologue(ctl, scheme);
}
// ------------------------------------------------------------------------
// sorted uniforms (U01SequenceSorted, u01.hpp:165-202):
//   lmax_n = fl(lmax_{n-1} + t_n),  t_n = log(1 - U_n) / (nn - n);  u_n = 1 - exp(lmax_n)
// vSMC accumulates lmax as a sequential FP64 left fold.  The terms are <= 0
// and round-to-nearest is symmetric, so -lmax_n is exactly the left fold of
// the non-negative x_n = -t_n: the passes below evaluate that fold with the
// exact-fold machinery (the same passes as the weight fold), so lmax is
// vSMC's bit for bit whenever the terms are.  The fold's apply pass turns the
// values into the uniforms, held as a running max per fold tile (the identity
// when exp is monotone), so the array is sorted by construction.
// ------------------------------------------------------------------------
// The terms x_n of n = n0 .. n0 + K - 1 (0 past nn).  Word n of the engine's
// stream is word (n + index) mod 4 of the block at counter ctr + (n + index)/4
// (the mapping of tf64_stream_word), so a run of consecutive n needs one
// Threefry4x64-20 block per four words.  n0 is a multiple of 4 in every
// caller, so the block changes at the same k in every lane.
template <int K>
__device__ __forceinline__ void sorted_terms(const Tf64State &e, uint64_t n0, uint64_t nn,
                                             double term[K])
{
    if (n0 >= nn) {  // past the sequence: no draws
#pragma unroll
        for (int k = 0; k < K; ++k) term[k] = 0.0;
        return;
    }
    u64x4 c0;
#pragma unroll
    for (int j = 0; j < 4; ++j) c0.v[j] = e.ctr[j];
    const uint64_t g0 = n0 + e.index;
    u64x4 b = {};
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t g = g0 + k, n = n0 + k;
        const unsigned w = (unsigned)(g & 3);
        if (k == 0 || w == 0) b = threefry4x64_20(ctr_add(c0, g >> 2), e.key);
        const uint64_t word = w == 0 ? b.v[0] : w == 1 ? b.v[1] : w == 2 ? b.v[2] : b.v[3];
        // x_n = -t_n, the negation of vSMC's rounded quotient
        term[k] = n < nn ? -(log(1 - canonical64(word)) / (double)(nn - n)) : 0.0;
    }
}

// Pass 1 of the uniform fold, over fold tiles of the nuni-element capacity:
// every x_n into ux (n < nuni; 0 from nn on), tile sums into tsum.  Each
// thread takes kFI consecutive n (whole Threefry blocks), so the sum of a tile
// is taken in another order than the weight fold's tile-sum pass; any fixed
// order is within the classification's error bound.
__global__ __launch_bounds__(kBlock) void k_sorted_terms(FilterCtl *ctl, double *ux, double *tsum,
                                                         uint64_t nuni)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kBlock / kWave];
    __shared__ double s_x[kFPad];
    const uint64_t nn = ctl->nn;
    const uint64_t base = blockIdx.x * (uint64_t)kFT;
    const int e0 = threadIdx.x * kFI;
    double term[kFI];
    sorted_terms<kFI>(ctl->eng, base + e0, nn, term);
    double s = 0;
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        s_x[fidx(e0 + k)] = term[k];
        s += term[k];
    }
    __syncthreads();
    // stored striped: one coalesced store instruction per element of each thread
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const int e = k * kBlock + threadIdx.x;
        if (base + e < nuni) ux[base + e] = s_x[fidx(e)];
    }
    s = block_reduce<kBlock>(s, AddOp(), s_red);
    if (threadIdx.x == 0) tsum[blockIdx.x] = s;
}

// single workgroup: exclusive running max of the tile maxima (fixed order)
__global__ __launch_bounds__(1024) void k_scan_tiles_max(FilterCtl *ctl, const double *tmax,
                                                         double *texcl, uint64_t ntiles)
{
    if (!ctl->resample) return;
    __shared__ double s_red[1024 / kWave];
    double carry = -1.0;
    for (uint64_t b = 0; b < ntiles; b += 1024) {
        const uint64_t i = b + threadIdx.x;
        const double v = i < ntiles ? tmax[i] : -1.0;
        double tot;
        const double incl = block_incl_scan<1024>(v, MaxOp(), s_red, &tot);
        // s_red holds per-wave inclusive maxima after the scan
        double ex = lane_up(incl, 1);
        const int lane = threadIdx.x & 63, wid = threadIdx.x / 64;
        if (lane == 0) ex = wid > 0 ? s_red[wid - 1] : -1.0;
        ex = ex < carry ? carry : ex;
        if (i < ntiles) texcl[i] = ex;
        carry = carry < tot ? tot : carry;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------
// Passes 1-4 give the fold total and the values at the specials; pass 5 runs
// when anything per element is wanted (inversion or fold values); passes 6-7
// when the ancestor layout is.
// lse_part (XWeight sources only): the move's tile partials of these
// log-weights, from which the tile starts follow without a tile-sum pass.
// join (optional): an event the inversion waits for (the sorted uniforms from
// the side stream)
template <typename XSrc, typename Inv>
static void launch_fold(FilterCtl *ctl, XSrc xs, Inv inv, uint64_t M, FoldScratch &fs,
                        const FoldOut &out, hipStream_t st, const double *lse_part = nullptr,
                        bool pre_started = false, Epi epi = Epi{nullptr, 0, 0},
                        hipEvent_t join = nullptr)
{
    fs.tiles = (M + kFT - 1) / kFT;
    const dim3 grid((unsigned)fs.tiles);
    if (lse_part && fs.wx) {
        if (!pre_started)  // else the reduce already wrote the starts
            hipLaunchKernelGGL(k_fold_scan_sum_lse, dim3(1), dim3(kScanBlock), 0, st, ctl,
                               lse_part, fs);
        // The systematic apply computes x from the log-weights again (an exp
        // per element is cheaper there than 8 more bytes per element, written
        // and read); the heavier inversions read the x the segment pass left.
        constexpr bool kCache = !std::is_same<Inv, InvSystematic>::value;
        hipLaunchKernelGGL((k_fold_seg<XSrc>), grid, dim3(kBlock), 0, st, ctl, xs, M, fs,
                           kCache ? fs.wx : (double *)nullptr);
        launch_scan_seg(ctl, fs, st);
        if (!Inv::kInv && !out.a_out) return;
        if (join) VSMCX_HIP(hipStreamWaitEvent(st, join, 0));
        if constexpr (kCache) {
            const XCached<XSrc> xc{fs.wx, xs};
            hipLaunchKernelGGL((k_fold_apply<XCached<XSrc>, Inv>), grid, dim3(kBlock), 0, st, ctl,
                               xc, inv, M, fs, out);
        } else {
            hipLaunchKernelGGL((k_fold_apply<XSrc, Inv>), grid, dim3(kBlock), 0, st, ctl, xs, inv,
                               M, fs, out);
        }
        if (!Inv::kInv || !out.anc) return;
        hipLaunchKernelGGL(k_fold_scan_cnt, dim3(1), dim3(kScanBlock), 0, st, ctl, fs, epi, M);
        hipLaunchKernelGGL(k_fold_layout, grid, dim3(kBlock), 0, st, ctl, fs, out, M);
        return;
    }
    hipLaunchKernelGGL((k_fold_tsum<XSrc>), grid, dim3(kBlock), 0, st, ctl, xs, M, fs);
    hipLaunchKernelGGL(k_fold_scan_sum, dim3(1), dim3(kScanBlock), 0, st, ctl, fs);
    if (fs.wx) {
        // later passes read the values the tile-sum pass wrote
        const XCached<XSrc> xc{fs.wx, xs};
        hipLaunchKernelGGL((k_fold_seg<XCached<XSrc>>), grid, dim3(kBlock), 0, st, ctl, xc, M, fs,
                           (double *)nullptr);
        launch_scan_seg(ctl, fs, st);
        if (!Inv::kInv && !out.a_out) return;
        if (join) VSMCX_HIP(hipStreamWaitEvent(st, join, 0));
        hipLaunchKernelGGL((k_fold_apply<XCached<XSrc>, Inv>), grid, dim3(kBlock), 0, st, ctl, xc,
                           inv, M, fs, out);
    } else {
        hipLaunchKernelGGL((k_fold_seg<XSrc>), grid, dim3(kBlock), 0, st, ctl, xs, M, fs,
                           (double *)nullptr);
        launch_scan_seg(ctl, fs, st);
        if (!Inv::kInv && !out.a_out) return;
        if (join) VSMCX_HIP(hipStreamWaitEvent(st, join, 0));
        hipLaunchKernelGGL((k_fold_apply<XSrc, Inv>), grid, dim3(kBlock), 0, st, ctl, xs, inv, M,
                           fs, out);
    }
    if (!Inv::kInv || !out.anc) return;
    hipLaunchKernelGGL(k_fold_scan_cnt, dim3(1), dim3(kScanBlock), 0, st, ctl, fs, epi, M);
    hipLaunchKernelGGL(k_fold_layout, grid, dim3(kBlock), 0, st, ctl, fs, out, M);
}

// Passes 3-7 once the tile starts are in fs.ts, x read from the source as is.
template <typename XSrc, typename Inv>
static void launch_fold_from_starts(FilterCtl *ctl, XSrc xs, Inv inv, uint64_t M,
                                    FoldScratch &fs, const FoldOut &out, hipStream_t st,
                                    Epi epi = Epi{nullptr, 0, 0}, hipEvent_t join = nullptr)
{
    const dim3 grid((unsigned)fs.tiles);
    hipLaunchKernelGGL((k_fold_seg<XSrc>), grid, dim3(kBlock), 0, st, ctl, xs, M, fs,
                       (double *)nullptr);
    launch_scan_seg(ctl, fs, st);
    if (!Inv::kInv && !out.a_out) return;
    if (join) VSMCX_HIP(hipStreamWaitEvent(st, join, 0));
    hipLaunchKernelGGL((k_fold_apply<XSrc, Inv>), grid, dim3(kBlock), 0, st, ctl, xs, inv, M, fs,
                       out);
    if (!Inv::kInv || !out.anc) return;
    hipLaunchKernelGGL(k_fold_scan_cnt, dim3(1), dim3(kScanBlock), 0, st, ctl, fs, epi, M);
    hipLaunchKernelGGL(k_fold_layout, grid, dim3(kBlock), 0, st, ctl, fs, out, M);
}

// The sorted uniforms (nn of them, nn <= rb.nuni): the exact fold of the x_n
// over the capacity (zeros from nn on leave the fold unchanged), in place in
// uscan, with its own tile sums and starts (the residual asum fold's tile sums
// in fs.tsum are still to be read by the scaled fold).  c: the control block
// the passes read (resample, nn, engine) and write (fold total, errors).
static void launch_uniform_fold(FilterCtl *c, ResampleBuffers &rb, hipStream_t st)
{
    FoldScratch &ufs = rb.ufs;
    const uint64_t nu = rb.nuni;
    const dim3 fgrid((unsigned)ufs.tiles);
    hipLaunchKernelGGL(k_sorted_terms, fgrid, dim3(kBlock), 0, st, c, rb.uscan, ufs.tsum, nu);
    hipLaunchKernelGGL(k_fold_scan_sum, dim3(1), dim3(kScanBlock), 0, st, c, ufs);
    // the apply pass leaves u = 1 - exp(-a) as a running max per fold tile
    launch_fold_from_starts(c, XArray{rb.uscan}, InvNone(), nu, ufs,
                            FoldOut{nullptr, nullptr, nullptr, nullptr, rb.uscan, rb.u_tmax}, st);
    hipLaunchKernelGGL(k_scan_tiles_max, dim3(1), dim3(1024), 0, st, c, rb.u_tmax, rb.u_tmaxex,
                       (uint64_t)ufs.tiles);
}

// The side stream's control block for one resample: the engine state and nn
// (multinomial: n; residual: n - R), resampling on (multinomial computes its
// uniforms before the step's decision; unused when it does not resample).
__global__ void k_side_begin(FilterCtl *uctl, const FilterCtl *ctl, int residual)
{
    if (threadIdx.x != 0) return;
    uctl->eng = ctl->eng;
    uctl->n = ctl->n;
    uctl->nn = residual ? ctl->n - ctl->r_int : ctl->n;
    uctl->resample = residual ? ctl->resample : 1;
    uctl->err = 0;
}

static void fork_uniforms(FilterCtl *ctl, ResampleBuffers &rb, hipStream_t st, int residual)
{
    SideFold &sd = rb.side;
    VSMCX_HIP(hipEventRecord(sd.fork, st));
    VSMCX_HIP(hipStreamWaitEvent(sd.s, sd.fork, 0));
    hipLaunchKernelGGL(k_side_begin, dim3(1), dim3(64), 0, sd.s, sd.uctl, ctl, residual);
    launch_uniform_fold(sd.uctl, rb, sd.s);
}

// Guesses of the inversion's uniform bounds for the weight fold's tiles: per
// tile t, F(lower end of the error interval around its approximate start
// ts[t]), by the 32-ary search of wave_first_greater2, two tiles per wave.
// The inversion verifies each guess (InvArray::wave_F2_guess), so these only
// spare it the search.
__global__ __launch_bounds__(kBlock) void k_side_windows(const FilterCtl *uctl, const double *ts,
                                                         uint64_t T, uint64_t M, InvArray inv,
                                                         uint32_t *guess)
{
    if (!uctl->resample) return;
    inv.bind(uctl);
    const uint64_t w = (blockIdx.x * (uint64_t)kBlock + threadIdx.x) / kWave;
    const uint64_t t0 = 2 * w, t1 = 2 * w + 1;
    if (t0 >= T) return;  // uniform over the wave
    double er, ea;
    fold_eps(M, &er, &ea);
    const double c0 = err_lo(ts[t0], er, ea);
    const double c1 = t1 < T ? err_lo(ts[t1], er, ea) : c0;
    uint64_t f0, f1;
    wave_first_greater2(inv, c0, c1, inv.nn, &f0, &f1);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        guess[t0] = (uint32_t)f0;
        if (t1 < T) guess[t1] = (uint32_t)f1;
    }
}

// The side stream's last step of a resample, and the join: with the weight
// fold's approximate tile starts in fs.ts (ts_ready), the guesses for its
// inversion after them.  Returns the guesses, or nullptr.
static const uint32_t *join_side(ResampleBuffers &rb, hipStream_t st, uint64_t M, bool ts_ready)
{
    SideFold &sd = rb.side;
    const uint32_t *g = nullptr;
    if (ts_ready && rb.uguess) {
        VSMCX_HIP(hipEventRecord(sd.starts, st));
        VSMCX_HIP(hipStreamWaitEvent(sd.s, sd.starts, 0));
        const uint64_t T = (M + kFT - 1) / kFT;
        const uint64_t waves = (T + 1) / 2;
        const unsigned blocks = (unsigned)((waves + kBlock / kWave - 1) / (kBlock / kWave));
        const InvArray inv{rb.uscan, rb.u_tmaxex, 0};
        hipLaunchKernelGGL(k_side_windows, dim3(blocks ? blocks : 1), dim3(kBlock), 0, sd.s,
                           sd.uctl, rb.fs.ts, T, M, inv, rb.uguess);
        g = rb.uguess;
    }
    VSMCX_HIP(hipEventRecord(sd.join, sd.s));
    return g;
}

void launch_side_uniforms(FilterCtl *ctl, ResampleBuffers &rb, hipStream_t st)
{
    if (!rb.side.s) return;
    fork_uniforms(ctl, rb, st, 0);
    rb.side.pending = true;
}

template <typename W>
static void resample_pipeline(FilterCtl *ctl, W w, uint64_t M, int scheme, ResampleBuffers &rb,
                              StepRow *hist, hipStream_t st, const double *lse_part = nullptr,
                              bool pre_started = false, bool encode = false)
{
    FoldOut out{rb.anc, rb.zl, rb.sl, rb.rep_out, rb.a_out};
    out.encode = encode && rb.anc;
    const bool residual = scheme == kResidual || scheme == kResidualStratified ||
                          scheme == kResidualSystematic;
    const bool sorted = scheme == kMultinomial || scheme == kResidual;
    // the uniforms on the side stream: multinomial's were forked at the start
    // of the step (launch_side_uniforms), residual's fork once R is known
    SideFold &sd = rb.side;
    const bool side = sorted && sd.s && (scheme == kResidual || sd.pending);
    sd.pending = false;
    FoldScratch &fs = rb.fs;
    fs.tiles = (M + kFT - 1) / kFT;
    if (residual) {
        // asum of the fractions: one exp per element, in the tile-sum pass
        hipLaunchKernelGGL((k_resid_tsum<W>), dim3((unsigned)fs.tiles), dim3(kBlock), 0, st, ctl,
                           w, M, fs);
        hipLaunchKernelGGL(k_fold_scan_sum_resid, dim3(1), dim3(kScanBlock), 0, st, ctl, fs);
        if (side) fork_uniforms(ctl, rb, st, 1);
        const XCached<XResid<W>> xr{fs.wx, XResid<W>{w, 0.0}};
        launch_fold_from_starts(ctl, xr, InvNone(), M, fs,
                                FoldOut{nullptr, nullptr, nullptr, nullptr, nullptr}, st);
        hipLaunchKernelGGL(k_residual_prologue, dim3(1), dim3(64), 0, st, ctl, scheme);
    } else if (!pre_started) {
        hipLaunchKernelGGL(k_plain_prologue, dim3(1), dim3(64), 0, st, ctl, scheme);
    }
    // with the ancestor lists the count scan carries the engine bookkeeping
    const int sys_like = (scheme == kSystematic || scheme == kResidualSystematic) ? 1 : 0;
    Epi epi{hist, rb.anc ? 1 : 0, sys_like};
    hipEvent_t join = nullptr;
    InvArray ia{rb.uscan, rb.u_tmaxex, 0};
    if (side) {
        join = sd.join;
        epi.uerr = &sd.uctl->err;
        // multinomial: the reduce wrote the tile starts (pre_started)
        if (!residual) ia.guess = join_side(rb, st, M, pre_started);
    } else if (sorted) {
        launch_uniform_fold(ctl, rb, st);
    }
    if (residual) {
        // the scaled fractions: tile starts from the asum fold's tile sums, x
        // and integer parts from what its tile-sum pass stored
        hipLaunchKernelGGL(k_fold_scan_sum_scaled, dim3(1), dim3(kScanBlock), 0, st, ctl, fs);
        if (side) ia.guess = join_side(rb, st, M, true);
        const XResidStored xs{fs.wx, fs.rep8, fs.rep32, 0.0};
        if (scheme == kResidual)
            launch_fold_from_starts(ctl, xs, ia, M, fs, out, st, epi, join);
        else if (scheme == kResidualStratified)
            launch_fold_from_starts(ctl, xs, InvStratified(), M, fs, out, st, epi);
        else
            launch_fold_from_starts(ctl, xs, InvSystematic(), M, fs, out, st, epi);
    } else {
        if (scheme == kMultinomial)
            launch_fold(ctl, w, ia, M, rb.fs, out, st, lse_part, pre_started, epi, join);
        else if (scheme == kStratified)
            launch_fold(ctl, w, InvStratified(), M, rb.fs, out, st, lse_part, pre_started, epi);
        else
            launch_fold(ctl, w, InvSystematic(), M, rb.fs, out, st, lse_part, pre_started, epi);
    }
    if (rb.anc && !out.encode) {
        const unsigned zg = (unsigned)((M + kTile - 1) / kTile) * 4;
        hipLaunchKernelGGL(k_zero_fill, dim3(zg ? zg : 1), dim3(kBlock), 0, st, ctl, rb.zl, rb.sl,
                           rb.anc, M);
    }
    if (!epi.on)
        hipLaunchKernelGGL(k_resample_epilogue, dim3(1), dim3(64), 0, st, ctl, hist, sys_like, M);
}

void launch_resample_from_log_weights(FilterCtl *ctl, const double *lw, uint64_t M, int scheme,
                                      ResampleBuffers &rb, StepRow *hist, hipStream_t st,
                                      const double *lse_part, bool pre_started, bool encode)
{
    resample_pipeline(ctl, XWeight{lw, 0.0, 0.0}, M, scheme, rb, hist, st, lse_part,
                      pre_started && lse_part, encode);
}

void launch_resample_from_weights(FilterCtl *ctl, const double *w, uint64_t M, int scheme,
                                  ResampleBuffers &rb, StepRow *hist, hipStream_t st)
{
    resample_pipeline(ctl, XArray{w}, M, scheme, rb, hist, st);
}

void launch_fold_values(FilterCtl *ctl, const double *x, uint64_t M, ResampleBuffers &rb,
                        hipStream_t st)
{
    launch_fold(ctl, XArray{x}, InvNone(), M, rb.fs,
                FoldOut{nullptr, nullptr, nullptr, nullptr, rb.a_out}, st);
}

__global__ __launch_bounds__(kBlock) void k_widen_u32(const uint32_t *in, uint64_t *out,
                                                      uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kBlock)
        out[i] = in[i];
}

void launch_widen_u32(const uint32_t *in, uint64_t *out, uint64_t n, hipStream_t st)
{
    if (!n) return;
    const uint64_t g = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_widen_u32, dim3((unsigned)(g < 2048 ? g : 2048)), dim3(kBlock), 0, st,
                       in, out, n);
}

}  // namespace vsmcx
