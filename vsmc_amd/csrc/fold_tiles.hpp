// fold_tiles.hpp -- the per-tile bodies of exact-fold resampling.
//
// vSMC's inversion (include/vsmc/resample/common.hpp:57-84) and cfrp_trans
// (include/vsmc/resample/common.hpp:87-118) as work on 2048-element tiles:
// tile sums, classification and segment summaries, the chain over special
// elements, exact values with inversion and replication, and the keep-in-place
// layout.  kernels_fold.hip runs them as grid passes over the tiles of one
// large particle system; kernels_batch.hip runs them tile after tile inside
// one workgroup per filter.  Both run the same code on the same data, so a
// filter's ancestors do not depend on which path resamples it.
#pragma once
#include <type_traits>

#include "device_util.hpp"
#include "fold_core.hpp"
#include "math_core.hpp"
#include "internal.hpp"

namespace vsmcx {

// Tiles of 2048 elements, 8 per thread.  Per-element data lives in LDS (one
// padding double per 8 keeps the thread-consecutive reads conflict-free) and
// the per-element loops are not unrolled, which keeps every instantiation
// small enough for the instruction cache.
constexpr int kFI = kFoldItems;
constexpr int kFT = kFoldTile;
constexpr int kFPad = kFT + kFT / kFI;
// Uniforms per window of the windowed inversions: the whole LDS array but one
// entry (stratified staging holds len + 1 draws), so a tile whose uniforms
// exceed its 2048 elements by up to 12% still takes one window.
constexpr int kWin = kFPad - 1;
__device__ __forceinline__ int fidx(int e) { return e + e / kFI; }

// ------------------------------------------------------------------------
// x sources: the sequence being folded
// ------------------------------------------------------------------------
struct XArray {  // given weights (compat entry, plain FP64 path)
    const double *w;
    __device__ void bind(const FilterCtl *) {}
    __device__ double x(uint64_t i) const { return w[i]; }
    __device__ uint64_t ipart(uint64_t) const { return 0; }
};

// The values of another source, read back from the array the tile-sum pass
// wrote (same values, one FP64 exp per element saved per later pass); the
// integer parts still come from the source.
template <typename XSrc>
struct XCached {
    const double *wx;
    XSrc src;
    __device__ void bind(const FilterCtl *c) { src.bind(c); }
    __device__ double x(uint64_t i) const { return wx[i]; }
    __device__ uint64_t ipart(uint64_t i) const { return src.ipart(i); }
};

struct XWeight {  // W_i = exp(lw_i - m) * coeff, computed from the log-weights
    const double *lw;
    double m, coeff;
    LwView lwv;  // lw in the filter's form
    __device__ void bind(const FilterCtl *c)
    {
        m = c->m;
        coeff = c->coeff;
        lwv = lw_view(c, lw);
    }
    __device__ double x(uint64_t i) const { return exp_lse(lwv(i) - m) * coeff; }
    __device__ uint64_t ipart(uint64_t) const { return 0; }
};

// residual.hpp:57-61: r_i = modf(N * W_i, &i_i)
template <typename W>
struct XResid {
    W w;
    double nd;
    __device__ void bind(const FilterCtl *c) { w.bind(c); nd = (double)c->n; }
    __device__ double x(uint64_t i) const
    {
        double ip;
        return modf(nd * w.x(i), &ip);
    }
    __device__ uint64_t ipart(uint64_t i) const
    {
        double ip;
        modf(nd * w.x(i), &ip);
        return (uint64_t)ip;
    }
};

// residual.hpp:62-63: r *= 1 / asum(r), from the fractions r_i and integer
// parts that the residual tile-sum pass stored (the values XResid computes),
// so the scaled fold takes no exp per element.
struct XResidStored {
    const double *r;
    const uint8_t *ip8;    // integer parts as bytes, 255: the part is in ip32
    const uint32_t *ip32;
    double rc;
    __device__ void bind(const FilterCtl *c) { rc = c->rcoeff; }
    __device__ double x(uint64_t i) const { return r[i] * rc; }
    __device__ uint64_t ipart(uint64_t i) const
    {
        const uint32_t b = ip8[i];
        return b < 255u ? b : ip32[i];
    }
};

// ------------------------------------------------------------------------
// uniform sources: F(c) = #{n < nn : v_n <= c} for the sorted sequence v
// ------------------------------------------------------------------------
struct InvNone {
    static constexpr bool kInv = false;
    static constexpr int kMinWaves = 1;  // waves per SIMD the apply kernel is compiled for
    static constexpr bool kWindow = false;
    static constexpr bool kWindowTmp = false;
    static constexpr bool kPairF = false;
    __device__ void bind(const FilterCtl *) {}
};

// Counts and indices of the uniform sequences stay below 2^31 (every entry
// point bounds N so), so their conversions go through 32 bits: exact, and one
// instruction each instead of a 64-bit sequence.
__device__ __forceinline__ double cnt_d(uint64_t n) { return (double)(uint32_t)n; }
// floor(est) clamped to [0, nn]
__device__ __forceinline__ uint64_t cnt_floor(double est, uint64_t nn)
{
    return !(est > 0) ? 0 : (est >= cnt_d(nn) ? nn : (uint64_t)(uint32_t)est);
}

// U01SequenceSystematic (include/vsmc/rng/u01.hpp:253-287): u_n = u0 + n*delta
struct InvSystematic {
    static constexpr bool kInv = true;
    static constexpr int kMinWaves = 6;  // waves per SIMD the apply kernel is compiled for
    static constexpr bool kWindow = false;
    static constexpr bool kWindowTmp = false;
    static constexpr bool kPairF = false;
    double u0, delta;
    uint64_t nn;
    __device__ void bind(const FilterCtl *c) { u0 = c->u0; delta = c->delta; nn = c->nn; }
    __device__ double v(uint64_t n) const { return u0 + cnt_d(n) * delta; }
    // [v(n) <= c] over n in [0, nn), 1 below the range and 0 above it: non-
    // increasing in n, and F(c) is its first zero.
    __device__ int le(int64_t n, double c) const
    {
        if (n < 0) return 1;
        if (n >= (int64_t)nn) return 0;
        return v((uint64_t)n) <= c;
    }
    // F(c) is floor((c - u0) nn) + 1 up to the roundings of that estimate
    // and of v (a few units of 2^-53 on values <= 1, with N < 2^31).  Nearly
    // always v(n0) <= c < v(n0 + 1) at the estimate n0, which fixes F =
    // n0 + 1 by monotonicity alone (two places); else F is counted over the
    // three places from n0 - 1 once the two places outside them confirm the
    // bracket; anything else takes the search.
    __device__ uint64_t F(double c) const
    {
        if (nn == 0) return 0;
        const uint64_t n0 = cnt_floor((c - u0) * cnt_d(nn), nn);
        const int64_t a = (int64_t)n0;
        const int at = le(a, c);
        if (at && !le(a + 1, c)) return n0 + 1;
        if (!at && le(a - 1, c)) return n0;
        if (le(a - 2, c) && !le(a + 2, c))
            return (uint64_t)(a - 1 + le(a - 1, c) + le(a, c) + le(a + 1, c));
        uint64_t n = n0;
        while (n < nn && v(n) <= c) ++n;
        while (n > 0 && v(n - 1) > c) --n;
        return n;
    }
};

// u_n of the stratified sequence, kept out of line (20 Threefry rounds).  The
// engine state is read where it lives, in the control block: a reference to
// a local copy would put that copy in scratch memory for the call.
__device__ __noinline__ double strat_u(const Tf64State *eng, double delta, uint64_t n)
{
    return canonical64(tf64_stream_word(*eng, n)) * delta + cnt_d(n) * delta;
}

// U01SequenceStratified (u01.hpp:212-243): u_n = U_n*delta + n*delta, with
// U_n the n-th draw of the resample engine; v_n = max(u_n, u_{n-1}) gives
// inversion's running-max semantics (u_{n-2} < u_n always).
struct InvStratified {
    static constexpr bool kInv = true;
    static constexpr int kMinWaves = 5;  // waves per SIMD the apply kernel is compiled for
    static constexpr bool kWindow = true;
    static constexpr bool kWindowTmp = true;  // fill_window keeps a table in L.tmp
    static constexpr bool kPairF = true;
    const Tf64State *eng;  // the control block's resample engine (global memory)
    double delta;
    uint64_t nn;
    __device__ void bind(const FilterCtl *c) { eng = &c->eng; delta = c->delta; nn = c->nn; }
    __device__ double u(uint64_t n) const { return strat_u(eng, delta, n); }
    __device__ double v(uint64_t n) const
    {
        const double a = u(n);
        if (n == 0) return a;
        const double b = u(n - 1);
        return a < b ? b : a;
    }
    // v(w0 .. w0+len-1) into s_v (len <= kWin).  The u_n come a Threefry block
    // (four consecutive draws) per thread; within a block v_n = max(u_n,
    // u_{n-1}) is formed in registers, and the first draw of each block takes
    // the previous block's last draw from s_last (kWinBlocks entries) after a
    // barrier.  Every s_v entry is written by one thread in each phase.
    __device__ void fill_window(double *s_v, double *s_last, uint64_t w0, int len) const
    {
        const uint64_t nlo = w0 ? w0 - 1 : 0, nhi = w0 + (uint64_t)len;
        const uint64_t index = eng->index;
        const uint64_t p0 = nlo + index, p1 = nhi + index;  // stream positions
        const uint64_t b0 = p0 / 4, b1 = (p1 - 1) / 4;       // blocks
        u64x4 c0;
        uint64_t key[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c0.v[j] = eng->ctr[j];
            key[j] = eng->key[j];
        }
        for (uint64_t bk = b0 + threadIdx.x; bk <= b1; bk += blockDim.x) {
            const u64x4 r = threefry4x64_20(ctr_add(c0, bk), key);
            double prev = -1.0;  // u of the block's previous draw in range
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint64_t p = 4 * bk + w;
                if (p >= p0 && p < p1) {
                    const uint64_t n = p - index;
                    const double u = canonical64(r.v[w]) * delta + cnt_d(n) * delta;
                    if (n >= w0) s_v[n - w0] = u < prev ? prev : u;
                    prev = u;
                }
            }
            s_last[bk - b0] = prev;
        }
        __syncthreads();
        // the first draw of every later block: 4 bk is in range for b0 < bk <= b1
        for (uint64_t bk = b0 + 1 + threadIdx.x; bk <= b1; bk += blockDim.x) {
            const uint64_t n = 4 * bk - index;
            if (n >= w0) {
                const uint64_t k = n - w0;
                const double b = s_last[bk - b0 - 1];
                s_v[k] = s_v[k] < b ? b : s_v[k];
            }
        }
    }
    // u_n = (n + U_n) delta with U_n in [0, 1), so the first n with v_n > c
    // lies within a few places of floor(c nn): a bracket [b0, b0 + 6) the
    // caller verifies at both ends before trusting it.
    __device__ bool bracket(double c, uint64_t *b0) const
    {
        const uint64_t e = cnt_floor(c * cnt_d(nn), nn);
        *b0 = e >= 2 ? e - 2 : 0;
        return true;
    }
    // F(c_lo) and F(c_hi) (first n with v_n > c, nn if none) in one pass of
    // the whole wave: lanes 0-31 take c_lo, lanes 32-63 c_hi.  Lane j of a
    // half draws u_m for m = b0 - 2 + j around the bracket of its value and
    // takes u_{m-1} from the lane below, so each v_m costs one Threefry block.
    // v is non-decreasing, so when v_{b0-1} <= c the first m with v_m > c
    // (m >= nn counts as above every value) is F(c).  Returns false when a
    // half cannot conclude that way; the caller then searches.
    __device__ bool wave_F2(double c_lo, double c_hi, uint64_t *f_lo, uint64_t *f_hi) const
    {
        const int lane = threadIdx.x & 63, j = lane & 31;
        const double c = lane < 32 ? c_lo : c_hi;
        uint64_t b0;
        bracket(c, &b0);
        const int64_t m = (int64_t)b0 - 2 + j;
        const bool in = m >= 0 && (uint64_t)m < nn;
        const double um = in ? u((uint64_t)m) : 0.0;
        const double up = __shfl_up(um, 1, 32);  // u_{m-1}, from lane j - 1
        bool gt = false;
        if (j >= 1 && m >= 0) {
            if ((uint64_t)m >= nn) {
                gt = true;
            } else {
                const double vm = m == 0 ? um : (um < up ? up : um);
                gt = vm > c;
            }
        }
        const uint64_t mask = __ballot(gt);
        const uint32_t hl = (uint32_t)mask, hh = (uint32_t)(mask >> 32);
        const uint64_t bl = __shfl(b0, 0), bh = __shfl(b0, 32);
        if (!hl || (hl & 2u) || !hh || (hh & 2u)) return false;
        *f_lo = bl - 2 + (uint64_t)(__ffs(hl) - 1);
        *f_hi = bh - 2 + (uint64_t)(__ffs(hh) - 1);
        return true;
    }
};

// A sorted uniform array held as a tile-wise running max over fold tiles: v_n =
// max(tile_prefix_max[n / kFoldTile], uscan[n]).
struct InvArray {
    static constexpr bool kInv = true;
    static constexpr int kMinWaves = 8;  // waves per SIMD the apply kernel is compiled for
    static constexpr bool kWindow = true;
    static constexpr bool kWindowTmp = false;  // fill_window reads the array
    static constexpr bool kPairF = false;
    const double *uscan;
    const double *tmax;  // exclusive running max of earlier tiles
    uint64_t nn;
    // optional: per fold tile t a guess of F(start of the tile), from the
    // tile's approximate start (k_side_windows); only a starting place, the
    // answer is verified where it is used
    const uint32_t *guess = nullptr;
    __device__ void bind(const FilterCtl *c) { nn = c->nn; }
    __device__ double v(uint64_t n) const
    {
        const double a = uscan[n], b = tmax[n / kFoldTile];
        return a < b ? b : a;
    }
    // F(c_lo) and F(c_hi) (first n with v_n > c, nn if none) from guesses:
    // lanes 0-31 take c_lo, lanes 32-63 c_hi, lane j of a half tests n =
    // g - 1 + j.  v is non-decreasing, so when v_{g-1} <= c (or g = 0) the
    // first n with v_n > c among the 32 places is F(c), whatever the guess.
    // Returns false (for the whole wave) when a half cannot conclude that way.
    __device__ bool wave_F2_guess(double c_lo, double c_hi, uint64_t g_lo, uint64_t g_hi,
                                  uint64_t *f_lo, uint64_t *f_hi) const
    {
        const int lane = threadIdx.x & 63, j = lane & 31;
        const double c = lane < 32 ? c_lo : c_hi;
        const int64_t n = (int64_t)(lane < 32 ? g_lo : g_hi) - 1 + j;
        const bool gt = n < 0 ? false : ((uint64_t)n >= nn ? true : v((uint64_t)n) > c);
        const uint64_t mask = __ballot(gt);
        const uint32_t hl = (uint32_t)mask, hh = (uint32_t)(mask >> 32);
        if (!hl || (hl & 1u) || !hh || (hh & 1u)) return false;
        *f_lo = g_lo - 1 + (uint64_t)(__ffs(hl) - 1);
        *f_hi = g_hi - 1 + (uint64_t)(__ffs(hh) - 1);
        return true;
    }
    __device__ void fill_window(double *s_v, double *, uint64_t w0, int len) const
    {
        for (int k = threadIdx.x; k < len; k += blockDim.x) s_v[k] = v(w0 + k);
    }
    __device__ bool bracket(double, uint64_t *) const { return false; }  // no closed form
};

// ------------------------------------------------------------------------
// the fold passes
// ------------------------------------------------------------------------
// Reduce-then-scan throughout: grid passes over 2048-element tiles and
// single-workgroup scans over the per-tile summaries.  No workgroup waits for
// another (a look-back chain across tiles would cross XCDs at every hop), and
// every prefix is formed in one fixed order, so the passes that re-derive the
// approximate prefix of an element get the same value bit for bit.
struct FoldOut {
    uint32_t *anc;       // [M] ancestor index (zero slots filled by k_zero_fill)
    uint32_t *zl;        // [M] zero-slot list
    uint32_t *sl;        // [M] surplus-copy list
    uint64_t *rep;       // [M] optional replication counts
    double *a_out;       // [M] optional exact fold values
    // The sorted-uniform fold (a_i = -lmax_i): with u_tmax set, a_out receives
    // u_i = 1 - exp(-a_i) held as a running max within the tile instead, and
    // u_tmax[tile] the tile's maximum.
    double *u_tmax = nullptr;
    // encoded ancestors: a zero slot's anc holds kAncZero | its zero rank, and
    // readers take the source from the surplus list (anc_source); no zero list
    // is written and no zero-fill pass runs
    bool encode = false;
};



constexpr int kScanBlock = 1024;  // single-workgroup tile scans
constexpr int kScanItems = 8;     // tiles per thread per round

struct SegOp {
    __device__ __forceinline__ Seg operator()(const Seg &a, const Seg &b) const { return seg_op(a, b); }
};
struct SatAddOp {
    __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return sat_add(a, b); }
};

// Exclusive segment prefix of each thread's run in its tile (identity for
// thread 0).  When every thread took the fast path in one binade b, the runs
// are tr_units(d_t, b), and tr_units(d, b) then tr_units(e, b) is
// tr_units(d + e, b) (saturating like tr_compose), so one integer scan of d
// gives the same prefixes as the transducer scan at a fraction of its cost.
// That is the case in every tile away from the few binade crossings of the
// running sum.
// uni_b (optional) receives the uniform binade, or kNoBinade when the tile
// took the general scan; with a uniform binade the prefix is tr_units(*ex_d,
// uni_b) (identity for thread 0), all the apply pass needs to store.
template <typename Lds>
__device__ __forceinline__ Seg thread_seg_prefix(const Seg &g, int fast_b, Lds &L,
                                                 int *uni_b = nullptr, uint64_t *ex_d = nullptr)
{
    if (threadIdx.x == 0) L.bu[0] = (uint64_t)(uint32_t)fast_b;
    __syncthreads();
    const int fb0 = (int)(uint32_t)L.bu[0];
    if (__syncthreads_and(fast_b == fb0 && fb0 != kNoBinade)) {
        const uint64_t ex =
            block_excl_scan<kBlock>(g.tr.d0, SatAddOp(), (uint64_t)0, L.u64, (uint64_t *)nullptr);
        Seg r = seg_identity();
        if (threadIdx.x != 0) r.tr = tr_units(ex, fb0);
        if (uni_b) {
            *uni_b = fb0;
            *ex_d = ex;
        }
        return r;
    }
    if (uni_b) {
        *uni_b = kNoBinade;
        *ex_d = 0;
    }
    return block_excl_scan<kBlock>(g, SegOp(), seg_identity(), L.seg, (Seg *)nullptr);
}

// The per-thread prefixes the segment pass leaves for the apply pass, per
// tile: thr[0..255] as Segs, or, when the whole tile took the fast path in
// one binade (fb[0] = that binade, else kNoBinade), only each thread's unit
// count as a uint64 in the first 2 KB of the tile's thr slots.
__device__ __forceinline__ void thr_store(Seg *thr, int32_t *fb, int uni_b, uint64_t ex_d,
                                          const Seg &pre)
{
    const int tid = threadIdx.x;
    if (tid == 0) fb[0] = uni_b;
    if (uni_b != kNoBinade)
        reinterpret_cast<uint64_t *>(thr)[tid] = ex_d;
    else
        thr[tid] = pre;
}
__device__ __forceinline__ Seg thr_load(const Seg *thr, const int32_t *fb)
{
    const int tid = threadIdx.x;
    const int b = fb[0];
    if (b == kNoBinade) return thr[tid];
    Seg r = seg_identity();
    if (tid != 0) r.tr = tr_units(reinterpret_cast<const uint64_t *>(thr)[tid], b);
    return r;
}

// F(c_lo) and F(c_hi) (first n in [0, nn) with v(n) > c, nn if none) by one
// wave: lanes 0-31 run a 32-ary search for c_lo while lanes 32-63 run one for
// c_hi, so the two searches share their rounds of dependent loads.
template <typename Inv>
__device__ void wave_first_greater2(const Inv &inv, double c_lo, double c_hi, uint64_t nn,
                                    uint64_t *f_lo, uint64_t *f_hi)
{
    const int lane = threadIdx.x & 63, j = lane & 31, sh = lane & 32;
    const double c = sh ? c_hi : c_lo;
    uint64_t lo = 0, hi = nn;  // the answer lies in [lo, hi]; the same in every lane of a half
    bool done = false;
    for (;;) {
        const bool act = !done && hi - lo > 32;
        if (!__ballot(act)) break;
        const uint64_t step = act ? (hi - lo + 31) / 32 : 0;
        const uint64_t n = lo + (uint64_t)j * step;
        const bool gt = act && (n < hi ? inv.v(n) > c : true);
        const uint32_t hm = (uint32_t)(__ballot(gt) >> sh);
        if (act) {
            const int g = hm ? __ffs(hm) - 1 : 32;
            if (g == 0) {
                hi = lo;
                done = true;
            } else {
                const uint64_t nlo = lo + (uint64_t)(g - 1) * step + 1;
                const uint64_t nhi = g < 32 ? lo + (uint64_t)g * step : hi;
                lo = nlo;
                hi = nhi < hi ? nhi : hi;
            }
        }
    }
    // at most 32 candidates left in a half that is not done
    const uint64_t n = lo + (uint64_t)j;
    const bool gt = done || (n < hi ? inv.v(n) > c : true);
    const uint32_t hm = (uint32_t)(__ballot(gt) >> sh);
    uint64_t r = lo;
    if (!done) {
        r = hm ? lo + (uint64_t)(__ffs(hm) - 1) : hi;
        r = r < hi ? r : hi;
    }
    *f_lo = __shfl(r, 0);
    *f_hi = __shfl(r, 32);
}

// x of one tile, striped (element k * kBlock + tid of the tile in v[k]); zero
// past the end, which classifies as negligible.  Every value is computed at an
// index clamped into [0, M) and zeroed past the end afterwards, so the loop
// has no branch and a thread's loads are issued together (a branch per
// element would wait out one load before the next).
template <typename XSrc>
__device__ __forceinline__ void tile_values(const XSrc &xs, uint64_t base, uint64_t M,
                                            double v[kFI])
{
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const uint64_t i = base + k * kBlock + threadIdx.x;
        const double t = xs.x(i < M ? i : M - 1);
        v[k] = i < M ? t : 0.0;
    }
}

// x of one tile into LDS (striped, coalesced)
template <typename XSrc>
__device__ __forceinline__ void load_tile(const XSrc &xs, uint64_t base, uint64_t M, double *s_x,
                                          double *wx_out = nullptr)
{
    double v[kFI];
    tile_values(xs, base, M, v);
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const int e = k * kBlock + threadIdx.x;
        const uint64_t i = base + e;
        s_x[fidx(e)] = v[k];
        if (wx_out && i < M) wx_out[i] = v[k];
    }
}

// The segment pass and the apply pass derive every element's approximate
// prefix from the tile's scanned start with the same code (the thread's run
// read into registers once, thread_sum, the block scan, the run's
// classification), on the same data, so they classify every element
// identically.
__device__ __forceinline__ void thread_run(const double *s_x, int e0, double xv[kFI])
{
#pragma unroll
    for (int k = 0; k < kFI; ++k) xv[k] = s_x[fidx(e0 + k)];
}
__device__ __forceinline__ double thread_sum(const double xv[kFI])
{
    double s = 0;
#pragma unroll
    for (int k = 0; k < kFI; ++k) s += xv[k];
    return s;
}
// the same sum read from LDS one value at a time (fewer registers live)
__device__ __forceinline__ double thread_sum_lds(const double *s_x, int e0)
{
    double s = 0;
#pragma unroll 1
    for (int k = 0; k < kFI; ++k) s += s_x[fidx(e0 + k)];
    return s;
}

// The run's segment summary: the fast path over the registers, the rare
// element-by-element path over the LDS copy (a loop that does not unroll
// would put a register array in scratch memory).
__device__ __forceinline__ Seg thread_run_seg(const double xv[kFI], const double *s_x, int e0,
                                              double s_thr, double er, double ea, int *fast_b)
{
    Seg g;
    if (run_seg_fast([&](int k) { return xv[k]; }, kFI, s_thr, er, ea, fast_b, &g)) return g;
    *fast_b = kNoBinade;
    return run_seg_elems([&](int k) { return s_x[fidx(e0 + k)]; }, kFI, s_thr, er, ea);
}

// Exclusive scan of n per-tile values by one workgroup of kScanBlock threads,
// each taking kScanItems consecutive tiles per round; fixed order.  The loads
// of a round are issued together (registers) so their latencies overlap.
struct NoVisit {
    template <typename T>
    __device__ void operator()(uint64_t, const T &) const {}
};

template <int BLOCK, typename T, typename Op, typename Visit = NoVisit, typename In = const T *>
__device__ T wg_scan_b(In in, T *out, uint64_t n, Op op, T ident, T *lds, Visit visit = Visit())
{
    T carry = ident;
    for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)BLOCK * kScanItems) {
        const uint64_t i0 = c0 + (uint64_t)threadIdx.x * kScanItems;
        T v[kScanItems];
        // loads at indices clamped into [0, n): no branch between them
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const T x = in[i0 + k < n ? i0 + k : n - 1];
            v[k] = i0 + k < n ? x : ident;
        }
        T loc = ident;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) loc = op(loc, v[k]);
        T tot;
        T run = op(carry, block_excl_scan<BLOCK>(loc, op, ident, lds, &tot));
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < n) {
                out[i0 + k] = run;
                visit(i0 + k, run);
            }
            run = op(run, v[k]);
        }
        carry = op(carry, tot);
    }
    return carry;
}

template <typename T, typename Op, typename Visit = NoVisit, typename In = const T *>
__device__ T wg_scan(In in, T *out, uint64_t n, Op op, T ident, T *lds, Visit visit = Visit())
{
    return wg_scan_b<kScanBlock, T, Op, Visit, In>(in, out, n, op, ident, lds, visit);
}

// Tile sums of the normalised weights x_i = exp(lw_i - m) coeff taken from the
// move's tile partials {max_t, sum_t exp(lw_i - max_t), ...} (5 doubles per
// tile, the move and fold tiles coincide): sum_t exp(max_t - m) coeff.  Each
// term differs from the x_i the fold adds by the roundings of the exp
// arguments, at most 2 u |lw_i - m| <= 1490 u relative for a term that does
// not underflow, plus a few u for the exps and products; fold_eps covers it.
struct TileSumsFromLse {
    const double *part;
    double m, coeff;
    __device__ double operator[](uint64_t t) const
    {
        const double mt = part[5 * t], se = part[5 * t + 1];  // both loads before the test
        return mt == -INFINITY ? 0.0 : se * exp_lse(mt - m) * coeff;
    }
};

// Tile sums of the scaled residual fractions r_i * rc from the tile sums of
// r_i: each differs from a sum of the fold's own products by the rounding of
// the products and of one more multiply, a few u relative, inside fold_eps.
// The same tile sums from the products se_t e^(M_t - m) that the reduction
// left in LDS (tile_partials_lse_block_stash): the same bits.
struct TileSumsStashed {
    const double *stash;
    double coeff;
    __device__ double operator[](uint64_t t) const { return stash[t] * coeff; }
};

struct TileSumsScaled {
    const double *tsum;
    double rc;
    __device__ double operator[](uint64_t t) const { return tsum[t] * rc; }
};

// ---- tile sum of x (any fixed order) -----------------------------------------
template <typename XSrc>
__device__ __forceinline__ double fold_tile_sum(const XSrc &xs, uint64_t base, uint64_t M,
                                                double *s_red, double *wx_out = nullptr)
{
    double v[kFI];
    tile_values(xs, base, M, v);
    double s = 0;
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const uint64_t i = base + k * kBlock + threadIdx.x;
        if (wx_out && i < M) wx_out[i] = v[k];
        s += v[k];
    }
    return block_reduce<kBlock>(s, AddOp(), s_red);
}

// LDS of the per-tile bodies; the table of Threefry blocks' last draws only
// where the inversion generates its uniforms (stratified).
constexpr int kWinBlocks = kWin / 4 + 3;  // Threefry blocks a window of kWin draws + 1 spans
template <bool kWindowBuf>
struct FoldLdsT {
    double x[kFPad];
    double tmp[kWindowBuf ? kWinBlocks : 1];
    double red[kBlock / kWave];
    Seg seg[kBlock / kWave];
    uint64_t u64[kBlock / kWave];
    uint64_t bu[2];
    double wend[kBlock / kWave];  // exact value at the end of each wave's runs
};

// ---- classification, tile segment summary, special records ----------------------
// The tile's specials go to spec_tile[k] in tile order, each with the run since
// the previous special of the tile (or since the tile start); the tile summary
// to *agg_out (written by the last thread).
template <typename XSrc, typename Lds>
__device__ void fold_tile_seg(const XSrc &xs, uint64_t base, uint64_t M, double ts,
                              SpecRec *spec_tile, Seg *agg_out, Lds &L, Seg *thr_out = nullptr,
                              int32_t *fb_out = nullptr, double *wx_out = nullptr)
{
    double *const s_x = L.x;
    __syncthreads();  // the LDS of a previous tile is no longer read
    const int tid = threadIdx.x;
    const int e0 = tid * kFI;
    load_tile(xs, base, M, s_x, wx_out);
    __syncthreads();
    double xv[kFI];
    thread_run(s_x, e0, xv);
    const double s_thr =
        ts + block_excl_scan<kBlock>(thread_sum(xv), AddOp(), 0.0, L.red, (double *)nullptr);
    double er, ea;
    fold_eps(M, &er, &ea);
    int fast_b;
    const Seg g = thread_run_seg(xv, s_x, e0, s_thr, er, ea, &fast_b);
    int uni_b;
    uint64_t ex_d;
    const Seg thr_pre = thread_seg_prefix(g, fast_b, L, &uni_b, &ex_d);
    // tile total from the last thread (a struct total through a pointer would
    // keep the struct in scratch memory)
    if (tid == kBlock - 1) *agg_out = seg_op(thr_pre, g);
    // the apply pass can start from these instead of classifying again
    if (thr_out) thr_store(thr_out, fb_out, uni_b, ex_d, thr_pre);
    if (g.cnt) {
        Seg pre = thr_pre;  // tile-local
        double s = s_thr;
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) {
            const double x = s_x[fidx(e0 + k)];
            const double sb = s;
            s = sb + x;
            const Seg el = elem_seg(x, sb, s, er, ea);
            if (el.cnt) {
                SpecRec *r = spec_tile + pre.cnt;
                r->run.d0 = pre.tr.d0;
                r->run.d1 = pre.tr.d1;
                r->run.pp = pre.tr.pp;
                r->run.b = pre.tr.b;
                r->x = x;
            }
            pre = seg_op(pre, el);
        }
    }
}

// ---- the chain over the specials in fold order -------------------------------------
// Thread 0's part: n records from LDS in fold order, one transducer
// application and one real FP64 add each; the next record is read while the
// current one is applied.
__device__ __forceinline__ void chain_run(const SpecRec *s_r, int n, double &a, double *s_v,
                                          int *bad)
{
    if (n <= 0) return;
    SpecRec cur = s_r[0];
#pragma unroll 1
    for (int k = 0; k < n; ++k) {
        const SpecRec nxt = s_r[k + 1 < n ? k + 1 : k];
        double v;
        if (!tr_apply(cur.run, a, &v)) {
            *bad = 1;
            v = a;
        }
        a = v + cur.x;
        s_v[k] = a;
        cur = nxt;
    }
}

// T tiles with exclusive segment summaries pre[] (their special counts also
// in cnt_lds when that is not null), records at spec[tile * kFT + k].  BLOCK
// threads take a chunk of specials each round: each finds its record (binary
// search over the counts), thread 0 runs the chain (one transducer application
// and one real FP64 add per special), and the values go to gval in fold
// order.  Returns the exact fold total in thread 0; *bad is set in thread 0
// when a transducer does not apply.
template <int BLOCK>
__device__ double fold_chain(const Seg *pre, const uint32_t *cnt_lds, uint64_t T, const Seg &tail,
                             const SpecRec *spec, double *gval, SpecRec *s_r, double *s_v,
                             int *bad)
{
    const int tid = threadIdx.x;
    const uint64_t S = tail.cnt;
    double a = 0.0;  // running exact value (thread 0)
    *bad = 0;
    for (uint64_t c0 = 0; c0 < S; c0 += BLOCK) {
        const uint64_t g = c0 + tid;
        if (g < S) {
            // last tile whose count of earlier specials is <= g
            uint64_t lo = 0, hi = T - 1;
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) / 2;
                const uint64_t c = cnt_lds ? cnt_lds[mid] : pre[mid].cnt;
                if (c <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const Seg p = pre[lo];
            const uint64_t k = g - p.cnt;
            SpecRec r = spec[lo * (uint64_t)kFT + k];
            if (k == 0) r.run = tr_compose(p.tr, r.run);
            s_r[tid] = r;
        }
        __syncthreads();
        if (tid == 0)
            chain_run(s_r, (int)(S - c0 < (uint64_t)BLOCK ? S - c0 : (uint64_t)BLOCK), a, s_v, bad);
        __syncthreads();
        if (g < S) gval[g] = s_v[tid];
        __syncthreads();
    }
    double tot = a;
    if (tid == 0 && !tr_apply(tail.tr, a, &tot)) {
        *bad = 1;
        tot = a;
    }
    return tot;
}

// ---- pass 5: exact values, inversion, replication --------------------------------
// Exact a_i = (run since the last special) applied to the value after that
// special.  Then c_i = F(a_i) (uniforms at or below a_i) and the replication
// counts as differences; per tile the number of zero slots and of surplus
// copies for the layout pass.
//
// Only the exact value before each thread's run comes from the segment
// machinery (the run's prefix transducer applied to the value after the last
// special); the run's own values are vSMC's fold steps a = fl(a + x) from that
// value, exact by construction.  Every run must then end where the next one
// starts (the tile's last run: at the next tile's start, from tpre_next when
// has_next, or at the fold total), which checks every start and so every value: a broken
// classification bound is reported, never returned.
template <typename XSrc, typename Inv, typename Lds>
__device__ void fold_tile_apply(FilterCtl *ctl, const XSrc &xs, const Inv &inv, uint64_t t,
                                uint64_t M, double ts, const Seg tpre, const Seg tpre_next,
                                bool has_next, const double *gval, const FoldOut &out,
                                uint32_t *rep32, uint8_t *rep8,
                                uint64_t *zs_out, Lds &L, const Seg *thr_in = nullptr,
                                const int32_t *fb_in = nullptr)
{
    // x, overwritten in place (each thread its own elements) by the exact
    // values; later the uniform window and the staged replication counts
    double *const s_x = L.x;
    double *const s_tmp = L.tmp;
    double *const s_a = s_x;
    double *const s_red = L.red;
    uint64_t *const s_u64 = L.u64;
    uint64_t *const s_bu = L.bu;
    __syncthreads();  // the LDS of a previous tile is no longer read

    const int tid = threadIdx.x;
    const uint64_t base = t * (uint64_t)kFT;
    const int e0 = tid * kFI;
    load_tile(xs, base, M, s_x);
    __syncthreads();

    // ---- exact fold values ----------------------------------------------------
    // The closed-form inversions take the thread's run in registers; the
    // sorted-array inversion and the plain fold read it from LDS a value at a
    // time, which keeps their register count (and waves per SIMD) at what
    // their latency-bound searches need.  Same values and sums either way.
    constexpr bool kReg = Inv::kInv && !(Inv::kWindow && !Inv::kWindowTmp);
    double xv[kFI];
    if constexpr (kReg) thread_run(s_x, e0, xv);
    Seg thr_pre;
    if (thr_in) {  // the segment pass's own classification of this thread's run
        thr_pre = thr_load(thr_in, fb_in);
    } else {
        const double tsum = kReg ? thread_sum(xv) : thread_sum_lds(s_x, e0);
        const double s_thr =
            ts + block_excl_scan<kBlock>(tsum, AddOp(), 0.0, s_red, (double *)nullptr);
        double er, ea;
        fold_eps(M, &er, &ea);
        int fast_b;
        Seg g;
        if constexpr (kReg)
            g = thread_run_seg(xv, s_x, e0, s_thr, er, ea, &fast_b);
        else
            g = run_seg([&](int k) { return s_x[fidx(e0 + k)]; }, kFI, s_thr, er, ea, &fast_b);
        thr_pre = thread_seg_prefix(g, fast_b, L);
    }
    const auto vb = [=](uint64_t c) { return gval[c]; };
    const RunStart rs = run_start(seg_op(tpre, thr_pre), vb);
    int bad = !rs.ok;
    double a = rs.a_prev;
    if constexpr (kReg) {
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            a = a + xv[k];
            s_a[fidx(e0 + k)] = a;
        }
    } else {
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) {
            const int e = fidx(e0 + k);
            a = a + s_x[e];
            s_a[e] = a;
        }
    }
    // exact a_{base-1}: thread 0's start, which only wave 0 reads below
    const double v_start = __shfl(rs.a_prev, 0);
    {
        // run boundaries: the previous lane's end, the previous wave's end,
        // and after the last run the next tile's start or the fold total
        const int lane = tid & (kWave - 1), wid = tid / kWave;
        const double up = __shfl_up(a, 1);
        if (lane == kWave - 1) L.wend[wid] = a;
        __syncthreads();
        double want = up;
        if (lane == 0) want = wid > 0 ? L.wend[wid - 1] : rs.a_prev;
        bool ok = same_value(rs.a_prev, want);
        if (tid == kBlock - 1) {
            double nx = ctl->fold_total;
            if (has_next &&
                !tr_apply(tpre_next.tr, tpre_next.cnt ? gval[tpre_next.cnt - 1] : 0.0, &nx))
                ok = false;
            ok = ok && same_value(a, nx);
        }
        if (!ok) bad = 1;
    }
    if (bad) atomicOr(&ctl->err, (int)kErrFoldClass);
    __syncthreads();

    if (out.a_out) {
        if (out.u_tmax) {
            // U01SequenceSorted's u = 1 - exp(lmax), lmax = -a (u01.hpp:185-187),
            // as inversion reads it: the running max (bins only advance)
            double mx = -1.0;
#pragma unroll 1
            for (int k = 0; k < kFI; ++k) {
                const double u = 1 - exp(-s_a[fidx(e0 + k)]);
                mx = mx < u ? u : mx;
                s_a[fidx(e0 + k)] = mx;
            }
            double tmax;
            const double mex = block_excl_scan<kBlock>(mx, MaxOp(), -1.0, s_red, &tmax);
#pragma unroll 1
            for (int k = 0; k < kFI; ++k) {
                const double u = s_a[fidx(e0 + k)];
                s_a[fidx(e0 + k)] = u < mex ? mex : u;
            }
            if (tid == 0) out.u_tmax[t] = tmax;
            __syncthreads();
        }
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) {
            const int e = k * kBlock + tid;
            const uint64_t i = base + e;
            if (i < M) out.a_out[i] = s_a[fidx(e)];
        }
    }

    if constexpr (!Inv::kInv) {
        return;
    } else {
        // ---- inversion: c_i = F(a_i) for i <= M-2 -------------------------------
        const uint64_t nn = inv.nn;
        const uint64_t last_cmp = M >= 2 ? M - 2 : 0;  // highest index compared
        uint32_t c[kFI];  // counts below 2^31
        double av[kFI];
#pragma unroll
        for (int k = 0; k < kFI; ++k) av[k] = s_a[fidx(e0 + k)];
        uint64_t cstart = 0;  // F(a_{base-1}); 0 for the first tile
        if constexpr (!Inv::kWindow) {
#pragma unroll
            for (int k = 0; k < kFI; ++k) {
                const uint64_t i = base + e0 + k;
                c[k] = (M >= 2 && i <= last_cmp) ? (uint32_t)inv.F(av[k]) : 0u;
            }
            if (t > 0 && tid == 0) cstart = inv.F(v_start);
        } else {
            // the tile holds uniforms [F(a_{base-1}), F(a_{last compared}))
            const bool need_hi = M >= 2 && base <= last_cmp;
            // the superset window: the stratified closed form puts F(v_start)
            // and F(c_hi) within [e - 1, e + 2] of e = floor(c nn), so [lo, hi)
            // from e(v_start) - 2 to e(c_hi) + 3 holds the tile's uniforms.
            // Checked once filled: v_lo <= v_start (every place below lo
            // counts for every element) and v_{hi-1} > c_hi (none from there
            // on does); otherwise the exact bounds are searched as below.
            bool superset = false;
            if constexpr (Inv::kPairF) {
                if (tid == 0) {
                    uint64_t mode = 0;
                    if (need_hi && nn > 0) {
                        const uint64_t il =
                            (base + kFT - 1 < last_cmp) ? base + kFT - 1 : last_cmp;
                        const double c_hi = s_a[fidx((int)(il - base))];
                        uint64_t lo, hi;
                        inv.bracket(v_start, &lo);
                        inv.bracket(c_hi, &hi);
                        if (t == 0) lo = 0;
                        hi = hi + 5 < nn ? hi + 5 : nn;
                        if (lo < hi && hi - lo <= (uint64_t)kWin) {
                            mode = 1;
                            s_bu[0] = lo;
                            s_bu[1] = hi;
                            L.wend[0] = v_start;
                            L.wend[1] = c_hi;
                        }
                    }
                    s_u64[0] = mode;
                }
                __syncthreads();
                if (s_u64[0]) {
                    const uint64_t lo = s_bu[0], hi = s_bu[1];
                    const int len = (int)(hi - lo);
                    inv.fill_window(s_x, s_tmp, lo, len);
                    __syncthreads();
                    superset = (lo == 0 || s_x[0] <= L.wend[0]) &&
                               (hi == nn || s_x[len - 1] > L.wend[1]);
                    if (!superset) __syncthreads();  // the window is read before the refill
                }
            }
            if (!superset && tid < kWave) {
                uint64_t lo = 0, hi = 0;
                if (M >= 2) {
                    const uint64_t il = (base + kFT - 1 < last_cmp) ? base + kFT - 1 : last_cmp;
                    // (c_hi saved before a superset window took the LDS)
                    const double c_hi = !need_hi ? v_start
                                        : (Inv::kPairF && s_u64[0]) ? L.wend[1]
                                                                    : s_a[fidx((int)(il - base))];
                    bool done = false;
                    // F(c_hi) >= F(v_start) = lo for sorted values; the max
                    // keeps the lower bound lo for any others
                    if constexpr (Inv::kPairF) {
                        uint64_t flo, fhi;
                        if (inv.wave_F2(v_start, c_hi, &flo, &fhi)) {
                            lo = t == 0 ? 0 : flo;
                            hi = need_hi ? (fhi < lo ? lo : fhi) : lo;
                            done = true;
                        }
                    }
                    if constexpr (std::is_same<Inv, InvArray>::value) {
                        // guesses of the bounds: the starts of this tile and
                        // of the next (the last tile has none for its end)
                        const uint64_t T = (M + kFT - 1) / kFT;
                        if (inv.guess && t + 1 < T) {
                            uint64_t flo, fhi;
                            if (inv.wave_F2_guess(v_start, c_hi, inv.guess[t], inv.guess[t + 1],
                                                  &flo, &fhi)) {
                                lo = t == 0 ? 0 : flo;
                                hi = need_hi ? (fhi < lo ? lo : fhi) : lo;
                                done = true;
                            }
                        }
                    }
                    if (!done) {
                        uint64_t flo, fhi;
                        wave_first_greater2(inv, v_start, c_hi, nn, &flo, &fhi);
                        lo = t == 0 ? 0 : flo;
                        hi = need_hi ? (fhi < lo ? lo : fhi) : lo;
                    }
                }
                if (tid == 0) {
                    s_bu[0] = lo;
                    s_bu[1] = hi;
                }
            }
            if (!superset) __syncthreads();
            const uint64_t wlo = s_bu[0], whi = s_bu[1];
            cstart = wlo;
#pragma unroll
            for (int k = 0; k < kFI; ++k) c[k] = (uint32_t)wlo;
            bool filled = superset;  // the superset window is already in LDS
            for (uint64_t w0 = wlo; w0 < whi; w0 += kWin) {
                const int len = (int)((whi - w0) < (uint64_t)kWin ? (whi - w0) : (uint64_t)kWin);
                if (!filled) {
                    __syncthreads();
                    inv.fill_window(s_x, s_tmp, w0, len);
                    __syncthreads();
                }
                filled = false;
                const double first = s_x[0], lastv = s_x[len - 1];
                // the thread's values are sorted: the first one by a binary
                // search over the window (the same rounds in every lane), each
                // later one from the previous position: four places in one
                // round (most values take fewer), then gallop and bisect
                // (first index with value > a)
                int pos = 0;
#pragma unroll
                for (int k = 0; k < kFI; ++k) {
                    const double a = av[k];
                    if (a < first) continue;
                    if (a >= lastv) {
                        c[k] += len;
                        pos = len;
                        continue;
                    }
                    if constexpr (Inv::kPairF) {
                        // u_n = (n + U_n) delta lies within a few units 2^-52 of
                        // [n, n + 1) / nn, so for b = floor(a nn) - w0 the window
                        // holds v_{b-2} <= a < v_{b+2}; both ends are checked, and
                        // the count is then b - 1 plus the places b-1, b, b+1 at
                        // or below a (v is non-decreasing), without a loop
                        const int b = (int)floor(a * cnt_d(nn) - cnt_d(w0));
                        if (b >= 2 && b + 2 < len) {
                            const double m2 = s_x[b - 2], m1 = s_x[b - 1], z0 = s_x[b],
                                         p1 = s_x[b + 1], p2 = s_x[b + 2];
                            if (m2 <= a && p2 > a) {
                                const int q = b - 1 + (m1 <= a) + (z0 <= a) + (p1 <= a);
                                c[k] += q;
                                pos = q;
                                continue;
                            }
                        }
                    }
                    int lo = pos, step = 1, hi = pos;
                    if (k == 0) {
                        hi = len - 1;  // v_{len-1} > a here
                    } else {
                        // places pos .. pos+3 (those past the window read as
                        // above a); v is non-decreasing, so the ones at or
                        // below a come first
                        const int j1 = pos + 1 < len ? pos + 1 : len - 1;
                        const int j2 = pos + 2 < len ? pos + 2 : len - 1;
                        const int j3 = pos + 3 < len ? pos + 3 : len - 1;
                        const double v0 = s_x[pos < len ? pos : len - 1], v1 = s_x[j1],
                                     v2 = s_x[j2], v3 = s_x[j3];
                        const int n = (pos < len && v0 <= a) + (pos + 1 < len && v1 <= a) +
                                      (pos + 2 < len && v2 <= a) + (pos + 3 < len && v3 <= a);
                        if (n < 4) {
                            c[k] += pos + n;
                            pos += n;
                            continue;
                        }
                        lo = hi = pos + 4;
                        while (hi < len && s_x[hi] <= a) {
                            lo = hi + 1;
                            hi += step;
                            step <<= 1;
                        }
                    }
                    if (hi > len) hi = len;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_x[mid] <= a)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    c[k] += lo;
                    pos = lo;
                }
            }
            if (superset && tid == 0) {
                // F(v_start): the places of the single window at or below it
                int q = 0;
                const int len = (int)(whi - wlo);
                while (q < len && s_x[q] <= L.wend[0]) ++q;
                cstart = t == 0 ? 0 : wlo + (uint64_t)q;
            }
#pragma unroll
            for (int k = 0; k < kFI; ++k) {
                const uint64_t i = base + e0 + k;
                if (!(M >= 2 && i <= last_cmp)) c[k] = 0;
            }
        }
        // c of the element before this thread's first item
        uint64_t cprev;
        {
            const int lane = tid & 63, wid = tid / 64;
            uint64_t up = lane_up(c[kFI - 1], 1);
            if (lane == 63) s_u64[wid] = c[kFI - 1];
            __syncthreads();
            if (lane == 0) up = wid > 0 ? s_u64[wid - 1] : 0;
            cprev = tid == 0 ? cstart : up;
            __syncthreads();
        }
        // replication counts, staged in LDS for striped stores
        uint32_t *s_rep = reinterpret_cast<uint32_t *>(s_x);
        uint64_t zs_thr = 0;  // zeros << 32 | surplus copies
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            const uint64_t i = base + e0 + k;
            uint64_t r = 0;
            if (i < M) {
                if (M == 1)
                    r = nn;
                else if (i <= last_cmp)
                    r = c[k] - cprev;
                else  // i == M-1
                    r = nn - cprev;
                r += xs.ipart(i);
                if (r == 0)
                    zs_thr += 1ull << 32;
                else if (r >= 2)
                    zs_thr += r - 1;
                if (out.rep) out.rep[i] = r;
            }
            cprev = c[k];
            s_rep[e0 + k] = (uint32_t)r;
            if (M >= 2 && i == last_cmp) ctl->n_final = c[k];
        }
        if (M == 1 && t == 0 && tid == 0) ctl->n_final = 0;
        if (!out.anc) return;  // replication only (M != N): no cfrp layout
        const uint64_t zs_tile = block_reduce<kBlock>(zs_thr, AddOp(), s_u64);
        if (tid == 0) *zs_out = zs_tile;
        // the counts as bytes for the layout pass, four elements a word (the
        // byte array spans whole tiles): a count of 255 or more goes in as
        // 255 with the count itself in rep32 (after every read of rep32 in
        // this tile: the residual integer parts, before the reduce above)
        uint32_t *const r8 = reinterpret_cast<uint32_t *>(rep8 + base);
#pragma unroll
        for (int w = 0; w < kFI / 4; ++w) {
            const int q = w * kBlock + tid;  // word q: elements 4q .. 4q+3
            uint32_t pk = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t r = s_rep[4 * q + j];
                pk |= (r < 255u ? r : 255u) << (8 * j);
                if (r >= 255u && base + 4 * q + j < M) rep32[base + 4 * q + j] = r;
            }
            r8[q] = pk;
        }
    }
}


// ---- pass 7: cfrp_trans layout (common.hpp:91-110) ------------------------------------
// anc[i] = i where rep_i > 0; the k-th zero slot (in index order) goes to the
// zero list at rank k, the k-th surplus copy (index order, rep_i - 1 copies
// of i) to the surplus list at rank k; k_zero_fill pairs them.
//
// The surplus list of a tile is written in one of two ways.  When it fits in
// LDS, every element with copies marks the tile-local rank of its first copy
// with its index, and an inclusive max-scan over the ranks carries each index
// over its copies (indices grow with rank); the list then goes out in striped
// stores.  Otherwise (more copies than LDS holds) each copy finds its element
// by a binary search over the running copy counts.  Both write the same list.

template <typename Lds>
__device__ void fold_tile_layout(FilterCtl *ctl, uint64_t t, uint64_t M, const uint32_t *rep32,
                                 const uint8_t *rep8, uint64_t zsx, const FoldOut &out, Lds &L)
{
    // kFPad entries each, element e at fidx(e): the per-thread loop reads
    // consecutive entries, which the padding spreads over the LDS banks.  Once
    // every thread holds its counts in registers, s_rep takes the tile's zero
    // slots in rank order instead (s_z), for striped stores of the zero list.
    uint32_t *const s_rep = reinterpret_cast<uint32_t *>(L.x);
    uint32_t *const s_z = s_rep;
    uint32_t *const s_sp = reinterpret_cast<uint32_t *>(L.x) + kFPad;
    uint64_t *const s_u64 = L.u64;
    __syncthreads();  // the LDS of a previous tile is no longer read
    const int tid = threadIdx.x;
    const uint64_t base = t * (uint64_t)kFT;
    const uint32_t b32 = (uint32_t)base;  // element indices stay below 2^32
    const int e0 = tid * kFI;
    const int nv = M - base < (uint64_t)kFT ? (int)(M - base) : kFT;  // elements of the tile
    // the apply pass's byte counts, four a word (255: the count is in rep32)
    const uint32_t *const r8 = reinterpret_cast<const uint32_t *>(rep8 + base);
#pragma unroll
    for (int w = 0; w < kFI / 4; ++w) {
        const int q = w * kBlock + tid;
        const uint32_t pk = r8[q];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = 4 * q + j;
            uint32_t r = (pk >> (8 * j)) & 255u;
            if (r == 255u && e < nv) r = rep32[base + e];
            s_rep[fidx(e)] = e < nv ? r : 1u;
        }
    }
    __syncthreads();
    uint32_t rv[kFI];
    uint64_t zs_thr = 0;
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const uint32_t r = s_rep[fidx(e0 + k)];
        rv[k] = r;
        if (r == 0)
            zs_thr += 1ull << 32;
        else if (r >= 2)
            zs_thr += r - 1;
    }
    uint64_t zs_tile;
    // (its barriers also end every read of s_rep before s_z is written)
    const uint64_t zs_excl =
        block_excl_scan<kBlock>(zs_thr, AddOp(), (uint64_t)0, s_u64, &zs_tile);
    const uint64_t z_base = zsx >> 32, s_base = zsx & 0xffffffffull;
    const uint32_t S_tile = (uint32_t)(zs_tile & 0xffffffffull);
    const bool direct = S_tile <= (uint32_t)kFPad;
    if (!out.encode) {
        uint32_t zl_local = (uint32_t)(zs_excl >> 32);  // zero slots before this thread (tile)
#pragma unroll
        for (int k = 0; k < kFI; ++k)
            if (e0 + k < nv && rv[k] == 0) s_z[zl_local++] = (uint32_t)(e0 + k);
    }
    if (!direct) {
        uint32_t srank_local = (uint32_t)(zs_excl & 0xffffffffull);  // surplus before this thread
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            if (rv[k] >= 2) srank_local += rv[k] - 1;
            s_sp[fidx(e0 + k)] = srank_local;
        }
    }
    if (out.encode) {
        // every slot's ancestor word, staged for striped stores: i, or the
        // zero marker with the slot's global zero rank (s_z is not used)
        __syncthreads();
        uint32_t zr = (uint32_t)(zs_excl >> 32);
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            const int e = e0 + k;
            uint32_t a = b32 + (uint32_t)e;
            if (rv[k] == 0) {
                const uint64_t rank = z_base + zr++;
                if (rank < M)
                    a = kAncZero | (uint32_t)rank;
                else  // never read through the surplus list out of bounds
                    atomicOr(&ctl->err, (int)kErrBounds);
            }
            s_z[fidx(e)] = a;
        }
        __syncthreads();
        uint32_t *const at = out.anc + base;
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            const int e = k * kBlock + tid;
            if (e < nv) at[e] = s_z[fidx(e)];
        }
    } else {
        uint32_t *const at = out.anc + base;
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            const int e = k * kBlock + tid;
            if (e < nv) at[e] = b32 + (uint32_t)e;
        }
        __syncthreads();
        // the zero list: ranks [z_base, z_base + Z_tile), striped
        const uint32_t Z_tile = (uint32_t)(zs_tile >> 32);
        for (uint32_t k = tid; k < Z_tile; k += kBlock) {
            if (z_base + k < M)
                out.zl[z_base + k] = b32 + s_z[k];
            else
                atomicOr(&ctl->err, (int)kErrBounds);
        }
    }
    // expand surplus copies: positions [s_base, s_base + S_tile)
    if (direct) {
        // s_sp is not otherwise used in this mode: marks, then the list
        for (uint32_t k = tid; k < S_tile; k += kBlock) s_sp[k] = 0;
        __syncthreads();
        uint32_t pos = (uint32_t)(zs_excl & 0xffffffffull);
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            if (rv[k] >= 2) {
                s_sp[pos] = (uint32_t)(e0 + k) + 1;  // 0: no mark
                pos += rv[k] - 1;
            }
        }
        __syncthreads();
        // inclusive max-scan of the marks: each thread a run of per entries
        constexpr int kPer = (kFPad + kBlock - 1) / kBlock;
        const uint32_t k0 = (uint32_t)tid * kPer;
        uint32_t mv[kPer];
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t v = k0 + j < S_tile ? s_sp[k0 + j] : 0u;
            run = v > run ? v : run;
            mv[j] = run;
        }
        const uint32_t carry = block_excl_scan<kBlock>(
            run, MaxOp(), 0u, reinterpret_cast<uint32_t *>(s_u64), (uint32_t *)nullptr);
#pragma unroll
        for (int j = 0; j < kPer; ++j)
            if (k0 + j < S_tile) s_sp[k0 + j] = mv[j] > carry ? mv[j] : carry;
        __syncthreads();
        for (uint32_t k = tid; k < S_tile; k += kBlock) {
            if (s_base + k < M)
                out.sl[s_base + k] = b32 + s_sp[k] - 1;
            else
                atomicOr(&ctl->err, (int)kErrBounds);
        }
        return;
    }
    for (uint32_t k = tid; k < S_tile; k += kBlock) {
        int lo = 0, hi = kFT - 1;  // first e with s_sp[e] > k
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_sp[fidx(mid)] > k)
                hi = mid;
            else
                lo = mid + 1;
        }
        if (s_base + k < M)
            out.sl[s_base + k] = b32 + (uint32_t)lo;
        else
            atomicOr(&ctl->err, (int)kErrBounds);
    }
}


// ---- resampling bookkeeping (one thread) ---------------------------------------------
// Non-residual schemes: nn = n uniforms, delta = 1/nn, and the systematic
// sequence's one constructor draw u0 = U * delta (rng/u01.hpp:253-263).
__device__ __forceinline__ void resample_prologue(FilterCtl *ctl, int scheme)
{
    ctl->nn = ctl->n;
    ctl->r_int = 0;
    ctl->delta = 1.0 / (double)ctl->nn;
    if (scheme == kSystematic) ctl->u0 = canonical64(tf64_stream_word(ctl->eng, 0)) * ctl->delta;
}

// Engine bookkeeping: vSMC's lazy sequences draw min(nn, n_final + 1)
// uniforms (rng/u01.hpp:38-40, resample/common.hpp:76-83); the systematic
// sequence draws one, in its constructor.
__device__ __forceinline__ void resample_epilogue(FilterCtl *ctl, StepRow *hist,
                                                  int systematic_like, uint64_t M)
{
    uint64_t draws;
    if (systematic_like)
        draws = 1;
    else
        draws = ctl->nn == 0 || M <= 1 ? 0
                : (ctl->n_final + 1 < ctl->nn ? ctl->n_final + 1 : ctl->nn);
    tf64_advance(ctl->eng, draws);
    ctl->draws = draws;
    ctl->prev_resampled = 1;
    if (hist) hist[ctl->t].draws = draws;
}

}  // namespace vsmcx
