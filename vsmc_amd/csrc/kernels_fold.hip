// kernels_fold.hip -- exact-fold resampling on the device.
//
// One kernel (k_fold_inv) per resample replaces vSMC's sequential
//   inversion  (include/vsmc/resample/common.hpp:57-84)  and
//   cfrp_trans (include/vsmc/resample/common.hpp:87-118),
// producing the replication counts, the keep-in-place ancestor layout and the
// lists the zero-slot fill (k_zero_fill) consumes.  The running weight sum is
// the exact left fold vSMC computes (see fold_core.hpp), so given the same
// weights and uniforms the ancestors are the reference's, at any N.
//
// Work decomposition: one 256-thread workgroup per 4096-particle tile, tiles
// claimed in dispatch order from an atomic counter, three decoupled
// look-backs (approximate FP64 prefix, exact fold value, zero/surplus counts).
#include "device_util.hpp"
#include "fold_core.hpp"
#include "internal.hpp"
#include "launch.hpp"

namespace vsmcx {

// ------------------------------------------------------------------------
// x sources: the sequence being folded
// ------------------------------------------------------------------------
struct XArray {  // given weights (compat entry, plain FP64 path)
    const double *w;
    __device__ void bind(const FilterCtl *) {}
    __device__ double x(uint64_t i) const { return w[i]; }
    __device__ uint64_t ipart(uint64_t) const { return 0; }
};

struct XWeight {  // W_i = exp(lw_i - m) * coeff, computed from the log-weights
    const double *lw;
    double m, coeff;
    __device__ void bind(const FilterCtl *c) { m = c->m; coeff = c->coeff; }
    __device__ double x(uint64_t i) const { return exp(lw[i] - m) * coeff; }
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

// residual.hpp:62-63: r *= 1 / asum(r)
template <typename W>
struct XResidScaled {
    XResid<W> r;
    double rc;
    __device__ void bind(const FilterCtl *c) { r.bind(c); rc = c->rcoeff; }
    __device__ double x(uint64_t i) const { return r.x(i) * rc; }
    __device__ uint64_t ipart(uint64_t i) const { return r.ipart(i); }
};

// ------------------------------------------------------------------------
// uniform sources: F(c) = #{n < nn : v_n <= c} for the sorted sequence v
// ------------------------------------------------------------------------
struct InvNone {
    static constexpr bool kInv = false;
    static constexpr bool kWindow = false;
    __device__ void bind(const FilterCtl *) {}
};

// U01SequenceSystematic (include/vsmc/rng/u01.hpp:253-287): u_n = u0 + n*delta
struct InvSystematic {
    static constexpr bool kInv = true;
    static constexpr bool kWindow = false;
    double u0, delta;
    uint64_t nn;
    __device__ void bind(const FilterCtl *c) { u0 = c->u0; delta = c->delta; nn = c->nn; }
    __device__ double v(uint64_t n) const { return u0 + (double)n * delta; }
    __device__ uint64_t F(double c) const
    {
        if (nn == 0) return 0;
        const double est = (c - u0) * (double)nn;
        uint64_t n = !(est > 0) ? 0 : (est >= (double)nn ? nn : (uint64_t)est);
        while (n < nn && v(n) <= c) ++n;
        while (n > 0 && v(n - 1) > c) --n;
        return n;
    }
};

// u_n of the stratified sequence, kept out of line (20 Threefry rounds)
__device__ __noinline__ double strat_u(const Tf64State &eng, double delta, uint64_t n)
{
    return canonical64(tf64_stream_word(eng, n)) * delta + (double)n * delta;
}

// U01SequenceStratified (u01.hpp:212-243): u_n = U_n*delta + n*delta, with
// U_n the n-th draw of the resample engine; v_n = max(u_n, u_{n-1}) gives
// inversion's running-max semantics (u_{n-2} < u_n always).
struct InvStratified {
    static constexpr bool kInv = true;
    static constexpr bool kWindow = true;
    Tf64State eng;
    double delta;
    uint64_t nn;
    __device__ void bind(const FilterCtl *c) { eng = c->eng; delta = c->delta; nn = c->nn; }
    __device__ double u(uint64_t n) const { return strat_u(eng, delta, n); }
    __device__ double v(uint64_t n) const
    {
        const double a = u(n);
        if (n == 0) return a;
        const double b = u(n - 1);
        return a < b ? b : a;
    }
    __device__ uint64_t point_F(double c) const
    {
        if (nn == 0) return 0;
        const double est = c * (double)nn;
        uint64_t n = !(est > 0) ? 0 : (est >= (double)nn ? nn : (uint64_t)est);
        while (n < nn && v(n) <= c) ++n;
        while (n > 0 && v(n - 1) > c) --n;
        return n;
    }
};

// A sorted uniform array held as a tile-wise running max: v_n =
// max(tile_prefix_max[n / kTile], uscan[n]).
struct InvArray {
    static constexpr bool kInv = true;
    static constexpr bool kWindow = true;
    const double *uscan;
    const double *tmax;  // exclusive running max of earlier tiles
    uint64_t nn;
    __device__ void bind(const FilterCtl *c) { nn = c->nn; }
    __device__ double v(uint64_t n) const
    {
        const double a = uscan[n], b = tmax[n / kTile];
        return a < b ? b : a;
    }
    __device__ uint64_t point_F(double c) const
    {
        // binary search for the first n with v_n > c
        uint64_t lo = 0, hi = nn;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (v(mid) <= c)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    }
};

// ------------------------------------------------------------------------
// decoupled look-back
// ------------------------------------------------------------------------
enum : uint64_t { kStNone = 0, kStAgg = 1, kStIncl = 2, kStNC = 3 };

__device__ __forceinline__ uint32_t lb_state(const uint64_t *flags, int64_t q, uint64_t epoch)
{
    const uint64_t f = atomic_load_rlx(flags + q);
    return (f >> 2) == epoch ? (uint32_t)(f & 3) : (uint32_t)kStNone;
}

__device__ __forceinline__ void lb_publish(uint64_t *flags, uint64_t q, uint64_t epoch,
                                           uint64_t st)
{
    atomic_store_rel(flags + q, (epoch << 2) | st);
}

// Exclusive prefix for tile t.  Called by all 64 lanes of one wave; every
// lane returns the result.  Tiles before 0 count as an inclusive `start`.
constexpr uint32_t kSpinLimit = 1u << 24;

template <typename A, typename V, typename LoadAgg, typename LoadIncl, typename Compose,
          typename Apply>
__device__ V lookback(uint64_t t, const uint64_t *flags, uint64_t epoch, V start, A ident,
                      LoadAgg load_agg, LoadIncl load_incl, Compose compose, Apply apply,
                      int *timed_out)
{
    const int lane = threadIdx.x & 63;
    A acc = ident;
    int64_t hi = (int64_t)t - 1;
    uint32_t spins = 0;
    while (true) {
        const int64_t q = hi - 63 + lane;
        uint32_t st = q < 0 ? (uint32_t)kStIncl : lb_state(flags, q, epoch);
        const uint64_t incl = __ballot(st == kStIncl);
        const int l_incl = incl ? 63 - __clzll(incl) : -1;
        const uint64_t notready = __ballot(lane > l_incl && st != kStAgg);
        if (notready) {
            // Wait on the nearest blocking predecessor alone (one broadcast load
            // per poll instead of 64), backing off, then rescan the window.
            // Bounded: a predecessor that never publishes is a bug; report it
            // and leave instead of hanging the device.
            const int wl = 63 - __clzll(notready);
            const int64_t qw = hi - 63 + wl;
            const uint32_t st_w = (uint32_t)__shfl((int)st, wl, 64);
            uint32_t backoff = 1;
            while (true) {
                const uint32_t s2 = lb_state(flags, qw, epoch);
                if (s2 != st_w) break;
                if (++spins > kSpinLimit) {
                    *timed_out = 1;
                    return start;
                }
                for (uint32_t k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(1);
                backoff = backoff < 8 ? backoff * 2 : 8;
            }
            continue;
        }
        fence_acq_agent();  // payloads published before the flags we saw
        A mine = (lane > l_incl) ? load_agg(q) : ident;
        A comp = wave_incl_scan(mine, compose);
        comp = lane_get(comp, 63);
        acc = compose(comp, acc);
        if (l_incl >= 0) {
            V base = start;
            const int64_t qi = hi - 63 + l_incl;
            if (qi >= 0) base = load_incl(qi);
            base = lane_get(base, l_incl);
            return apply(acc, base);
        }
        hi -= 64;
    }
}

// ------------------------------------------------------------------------
// the kernel
// ------------------------------------------------------------------------
struct FoldOut {
    uint32_t *anc;       // [M] ancestor index (zero slots filled by k_zero_fill)
    uint32_t *zl;        // [M] zero-slot list
    uint32_t *sl;        // [M] surplus-copy list
    uint64_t *rep;       // [M] optional replication counts
    double *a_out;       // [M] optional exact fold values
};

// Tiles of 2048 elements, 8 per thread.  Per-element data lives in LDS (one
// padding double per 8 keeps the thread-consecutive reads conflict-free) and
// the per-element loops are not unrolled, which keeps every instantiation
// small enough for the instruction cache.
constexpr int kFI = kFoldItems;
constexpr int kFT = kFoldTile;
constexpr int kFPad = kFT + kFT / kFI;
__device__ __forceinline__ int fidx(int e) { return e + e / kFI; }

// First n in [lo, hi) with v(n) > c (hi if none), v non-decreasing.  All 64
// lanes of a wave probe 64 evenly spaced points per round (64-ary search).
template <typename Inv>
__device__ uint64_t wave_first_greater(const Inv &inv, double c, uint64_t lo, uint64_t hi)
{
    const int lane = threadIdx.x & 63;
    while (hi - lo > 64) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t n = lo + (uint64_t)lane * step;
        const bool gt = n < hi ? inv.v(n) > c : true;
        const uint64_t mask = __ballot(gt);
        const int g = mask ? __ffsll((unsigned long long)mask) - 1 : 64;
        if (g == 0) return lo;
        const uint64_t nlo = lo + (uint64_t)(g - 1) * step + 1;
        const uint64_t nhi = g < 64 ? lo + (uint64_t)g * step : hi;
        lo = nlo;
        hi = nhi < hi ? nhi : hi;
    }
    const uint64_t n = lo + (uint64_t)lane;
    const bool gt = n < hi ? inv.v(n) > c : true;
    const uint64_t mask = __ballot(gt);
    if (mask == 0) return hi;  // all 64 remaining candidates are <= c
    const uint64_t r = lo + (uint64_t)(__ffsll((unsigned long long)mask) - 1);
    return r < hi ? r : hi;
}

template <typename XSrc, typename Inv>
__global__ __launch_bounds__(kBlock) void k_fold_inv(FilterCtl *ctl, XSrc xs, Inv inv, uint64_t M,
                                                     FoldScratch fs, FoldOut out)
{
    if (!ctl->resample) return;
    __shared__ double s_x[kFPad];  // x; later the uniform window / surplus prefix
    __shared__ double s_a[kFPad];  // exact fold values
    __shared__ double s_red[kBlock / kWave];
    __shared__ Tr s_tr[kBlock / kWave];
    __shared__ uint64_t s_u64[kBlock / kWave];
    __shared__ int s_int[kBlock / kWave];
    __shared__ unsigned long long s_tile;
    __shared__ double s_bd[2];
    __shared__ uint64_t s_bu[4];
    __shared__ int s_timeout;

    const int tid = threadIdx.x;
    if (tid == 0) {
        s_tile = atomicAdd(fs.tile_counter, 1ull);
        s_timeout = 0;
    }
    const uint64_t epoch = *fs.epoch;
    xs.bind(ctl);
    inv.bind(ctl);
    __syncthreads();
    const uint64_t t = s_tile;
    const uint64_t base = t * (uint64_t)kFT;
    const int e0 = tid * kFI;  // this thread's elements: e0 .. e0 + kFI - 1

    // ---- load x (striped, coalesced) into LDS --------------------------------
#pragma unroll
    for (int k = 0; k < kFI; ++k) {
        const int e = k * kBlock + tid;
        const uint64_t i = base + e;
        s_x[fidx(e)] = i < M ? xs.x(i) : 0.0;
    }
    __syncthreads();

    // ---- approximate prefix: tile sum, look-back ------------------------------
    double tsum = 0;
#pragma unroll 1
    for (int k = 0; k < kFI; ++k) tsum += s_x[fidx(e0 + k)];
    double tile_sum;
    block_incl_scan<kBlock>(tsum, AddOp(), s_red, &tile_sum);
    if (tid == 0) {
        if (t == 0)
            fs.incl_sum[t] = tile_sum;
        else
            fs.agg_sum[t] = tile_sum;
        lb_publish(fs.flags_sum, t, epoch, t == 0 ? kStIncl : kStAgg);
    }
    if (tid < kWave) {
        double S = 0.0;
        if (t > 0) {
            const double *agg = fs.agg_sum, *inc = fs.incl_sum;
            S = lookback<double, double>(
                t, fs.flags_sum, epoch, 0.0, 0.0, [=](int64_t q) { return agg[q]; },
                [=](int64_t q) { return inc[q]; }, [](double a, double b) { return a + b; },
                [](double a, double b) { return a + b; }, &s_timeout);
        }
        if (tid == 0) {
            s_bd[0] = S;
            if (t > 0) {
                fs.incl_sum[t] = S + tile_sum;
                lb_publish(fs.flags_sum, t, epoch, kStIncl);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && s_timeout) atomicOr(&ctl->err, (int)kErrTimeout);
    const double S_t = s_bd[0];

    // ---- classify the tile --------------------------------------------------
    double eps_rel, eps_abs;
    fold_eps(M, &eps_rel, &eps_abs);
    const int b0 = certain_binade(S_t, eps_rel, eps_abs);
    const int b1 = certain_binade(S_t + tile_sum, eps_rel, eps_abs);
    const bool clean = t > 0 && b0 == b1 && b0 != kNoBinade;
    const auto compose = [](const Tr &a, const Tr &c) { return tr_compose(a, c); };

    // ---- per-thread transducers, computed before the exact start is known ------
    // Clean tiles: binade bA = b0; publish the tile transducer.  Other tiles:
    // bA = lowest binade the start can be in, bB = bA + 1; prefix scan in bA and
    // reverse (suffix) scan in bB let the end value follow from the exact start
    // with one walk and one real add when the tile holds a single crossing.
    const int bA = clean ? b0 : binade_of(S_t - (S_t * eps_rel + eps_abs) > 0
                                              ? S_t - (S_t * eps_rel + eps_abs) : 0.0);
    const int bB = bA + 1;
    Tr PA = tr_identity(), TA = tr_identity(), RB = tr_identity(), TB = tr_identity();
    if (t > 0) {
        Tr thr = tr_identity();
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) thr = tr_compose(thr, tr_element(s_x[fidx(e0 + k)], bA));
        PA = block_excl_scan<kBlock>(thr, compose, tr_identity(), s_tr, &TA);
        if (clean) {
            if (tid == 0) {
                fs.agg_d0[t] = TA.d0;
                fs.agg_d1[t] = TA.d1;
                fs.agg_pp[t] = TA.pp;
                fs.agg_b[t] = TA.b;
                lb_publish(fs.flags_fold, t, epoch, kStAgg);
            }
        } else {
            if (tid == 0) lb_publish(fs.flags_fold, t, epoch, kStNC);
            Tr thb = tr_identity();
#pragma unroll 1
            for (int k = 0; k < kFI; ++k) thb = tr_compose(thb, tr_element(s_x[fidx(e0 + k)], bB));
            // reverse inclusive scan: z_k = T_{255-k} o ... o T_255, via reversed order
            Tr *s_rev = reinterpret_cast<Tr *>(s_a);  // kBlock entries, s_a not yet in use
            s_rev[kBlock - 1 - tid] = thb;
            __syncthreads();
            const Tr y = s_rev[tid];
            __syncthreads();
            const Tr z = block_incl_scan<kBlock>(
                y, [](const Tr &p, const Tr &q) { return tr_compose(q, p); }, s_tr, &TB);
            s_rev[tid] = z;
            __syncthreads();
            RB = tid < kBlock - 1 ? s_rev[kBlock - 2 - tid] : tr_identity();  // threads > tid
            __syncthreads();
        }
    }

    // ---- exact start value of the tile -----------------------------------------
    if (tid < kWave) {
        double v0 = 0.0;
        if (t > 0) {
            const uint64_t *ad0 = fs.agg_d0, *ad1 = fs.agg_d1;
            const uint32_t *app = fs.agg_pp;
            const int32_t *ab = fs.agg_b;
            const double *vf = fs.val_fold;
            int *errp = &ctl->err;
            v0 = lookback<Tr, double>(
                t, fs.flags_fold, epoch, 0.0, tr_identity(),
                [=](int64_t q) {
                    Tr a;
                    a.d0 = ad0[q];
                    a.d1 = ad1[q];
                    a.pp = app[q];
                    a.b = ab[q];
                    return a;
                },
                [=](int64_t q) { return vf[q]; }, compose,
                [=](const Tr &a, double v) {
                    double r;
                    if (!tr_apply(a, v, &r)) {
                        atomicOr(errp, (int)kErrFoldClass);
                        r = v;
                    }
                    return r;
                },
                &s_timeout);
        }
        if (tid == 0) {
            s_bd[1] = v0;
            if (clean) {  // publish the inclusive value right away
                double e;
                if (!tr_apply(TA, v0, &e)) atomicOr(&ctl->err, (int)kErrFoldClass);
                fs.val_fold[t] = e;
                lb_publish(fs.flags_fold, t, epoch, kStIncl);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && s_timeout) atomicOr(&ctl->err, (int)kErrTimeout);
    const double v0 = s_bd[1];
    bool published = clean;

    if (clean) {
        // every element: the thread prefix applied to the start, then a walk
        double v;
        if (!tr_apply(PA, v0, &v)) {
            atomicOr(&ctl->err, (int)kErrFoldClass);
            v = v0;
        }
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) {
            const int e = e0 + k;
            double v2;
            if (!tr_apply(tr_element(s_x[fidx(e)], bA), v, &v2)) {
                atomicOr(&ctl->err, (int)kErrFoldClass);
                v2 = v;
            }
            s_a[fidx(e)] = v2;
            v = v2;
        }
        __syncthreads();
    } else if (t > 0) {
        // ---- end value first (the next tiles wait for it) ---------------------
        const int b = binade_of(v0);
        int ok = 0;
        double endv = 0;
        if (b == bB) {
            ok = tr_apply(TB, v0, &endv);
        } else if (b == bA) {
            int first_fail = kFT;
            double vprev = 0, vt;
            if (!tr_apply(PA, v0, &vt)) {
                first_fail = e0;
            } else {
#pragma unroll 1
                for (int k = 0; k < kFI; ++k) {
                    double v2;
                    if (!tr_apply(tr_element(s_x[fidx(e0 + k)], bA), vt, &v2)) {
                        first_fail = e0 + k;
                        vprev = vt;
                        break;
                    }
                    vt = v2;
                }
            }
            const int j = block_reduce<kBlock>(first_fail, MinOp(), s_int);
            if (j >= kFT) {
                ok = tr_apply(TA, v0, &endv);
            } else {
                if (tid == j / kFI) {
                    // owner: one real add at the crossing, then the rest in bB
                    const double aj = vprev + s_x[fidx(j)];
                    Tr part = tr_identity();
#pragma unroll 1
                    for (int e = j + 1; e < e0 + kFI; ++e)
                        part = tr_compose(part, tr_element(s_x[fidx(e)], bB));
                    double w;
                    int okk = binade_of(aj) == bB && tr_apply(part, aj, &w);
                    double en = 0;
                    if (okk) okk = tr_apply(RB, w, &en);
                    s_bd[0] = en;
                    s_int[0] = okk;
                }
                __syncthreads();
                ok = s_int[0];
                endv = s_bd[0];
                __syncthreads();
            }
        }
        if (ok && tid == 0) {
            fs.val_fold[t] = endv;
            lb_publish(fs.flags_fold, t, epoch, kStIncl);
        }
        published = ok;
        __syncthreads();
    }

    if (!clean) {
        // ---- exact values of every element: carry the fold through the tile ----
        // Each pass composes the elements from `pos` in the binade of the running
        // value; the first element whose result would leave that binade gets one
        // real FP64 add, and the next pass starts after it.  The first tile
        // starts at 0 and crosses a binade every doubling, so its first
        // elements are folded one by one.
        double cur = v0;
        int pos = 0;
        if (t == 0) {
            constexpr int kSerial = 256;
            if (tid == 0) {
                double a = 0;
                for (int e = 0; e < kSerial; ++e) {
                    a += s_x[fidx(e)];
                    s_a[fidx(e)] = a;
                }
            }
            __syncthreads();
            cur = s_a[fidx(kSerial - 1)];
            pos = kSerial;
        }
        while (pos < kFT) {
            const int b = binade_of(cur);
            Tr thr = tr_identity();
#pragma unroll 1
            for (int k = 0; k < kFI; ++k) {
                const int e = e0 + k;
                if (e >= pos) thr = tr_compose(thr, tr_element(s_x[fidx(e)], b));
            }
            const Tr excl =
                block_excl_scan<kBlock>(thr, compose, tr_identity(), s_tr, (Tr *)nullptr);
            int first_fail = kFT;
            double v;
            if (!tr_apply(excl, cur, &v)) {
                first_fail = e0;
            } else {
#pragma unroll 1
                for (int k = 0; k < kFI; ++k) {
                    const int e = e0 + k;
                    if (e < pos) continue;
                    double v2;
                    if (!tr_apply(tr_element(s_x[fidx(e)], b), v, &v2)) {
                        first_fail = e;
                        break;
                    }
                    s_a[fidx(e)] = v2;
                    v = v2;
                }
            }
            const int j = block_reduce<kBlock>(first_fail, MinOp(), s_int);
            if (j >= kFT) break;
            if (tid == j / kFI) {
                const double prev = (j == pos) ? cur : s_a[fidx(j - 1)];
                s_a[fidx(j)] = prev + s_x[fidx(j)];
            }
            __syncthreads();
            cur = s_a[fidx(j)];
            pos = j + 1;
            __syncthreads();
        }
        __syncthreads();
        if (tid == 0) {
            if (!published) {
                fs.val_fold[t] = s_a[fidx(kFT - 1)];
                lb_publish(fs.flags_fold, t, epoch, kStIncl);
            } else if (fs.val_fold[t] != s_a[fidx(kFT - 1)]) {
                atomicOr(&ctl->err, (int)kErrFoldClass);
            }
        }
    }
    const double v_start = v0;
    if (out.a_out) {
#pragma unroll 1
        for (int k = 0; k < kFI; ++k) {
            const int e = k * kBlock + tid;
            const uint64_t i = base + e;
            if (i < M) out.a_out[i] = s_a[fidx(e)];
        }
    }
    if (t == fs.tiles - 1 && tid == 0) ctl->fold_total = s_a[fidx((int)((M - 1) - base))];

    if constexpr (!Inv::kInv) {
        return;
    } else {
        // ---- inversion: c_i = F(a_i) for i <= M-2 -------------------------------
        const uint64_t nn = inv.nn;
        const uint64_t last_cmp = M >= 2 ? M - 2 : 0;  // highest index compared
        uint64_t c[kFI];
        double av[kFI];
#pragma unroll
        for (int k = 0; k < kFI; ++k) av[k] = s_a[fidx(e0 + k)];
        uint64_t cstart = 0;  // F(a_{base-1}); 0 for the first tile
        if constexpr (!Inv::kWindow) {
#pragma unroll
            for (int k = 0; k < kFI; ++k) {
                const uint64_t i = base + e0 + k;
                c[k] = (M >= 2 && i <= last_cmp) ? inv.F(av[k]) : 0;
            }
            if (t > 0) cstart = inv.F(v_start);
        } else {
            // the tile holds uniforms [F(a_{base-1}), F(a_{last compared}))
            if (tid < kWave) {
                uint64_t lo = 0, hi = 0;
                if (M >= 2) {
                    lo = t == 0 ? 0 : wave_first_greater(inv, v_start, 0, nn);
                    hi = lo;
                    if (base <= last_cmp) {
                        const uint64_t il =
                            (base + kFT - 1 < last_cmp) ? base + kFT - 1 : last_cmp;
                        hi = wave_first_greater(inv, s_a[fidx((int)(il - base))], lo, nn);
                    }
                }
                if (tid == 0) {
                    s_bu[0] = lo;
                    s_bu[1] = hi;
                }
            }
            __syncthreads();
            const uint64_t wlo = s_bu[0], whi = s_bu[1];
            cstart = wlo;
#pragma unroll
            for (int k = 0; k < kFI; ++k) c[k] = wlo;
            for (uint64_t w0 = wlo; w0 < whi; w0 += kFPad) {
                const int len = (int)((whi - w0) < (uint64_t)kFPad ? (whi - w0) : (uint64_t)kFPad);
                __syncthreads();
#pragma unroll 1
                for (int k = tid; k < len; k += kBlock) s_x[k] = inv.v(w0 + k);
                __syncthreads();
                const double first = s_x[0], lastv = s_x[len - 1];
#pragma unroll
                for (int k = 0; k < kFI; ++k) {
                    const double a = av[k];
                    if (a < first) continue;
                    if (a >= lastv) {
                        c[k] += len;
                        continue;
                    }
                    int lo = 0, hi = len;  // first index with value > a
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_x[mid] <= a)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    c[k] += lo;
                }
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
        // replication counts
        uint32_t rep[kFI];
        uint64_t zs_thr = 0;  // zeros << 32 | surplus
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
            }
            cprev = c[k];
            rep[k] = (uint32_t)r;
            if (i < M) {
                if (r == 0)
                    zs_thr += 1ull << 32;
                else if (r >= 2)
                    zs_thr += r - 1;
            }
            if (out.rep && i < M) out.rep[i] = r;
            if (M >= 2 && i == last_cmp) ctl->n_final = c[k];
        }
        if (M == 1 && t == 0 && tid == 0) ctl->n_final = 0;
        if (!out.anc) return;  // replication only (M != N): no cfrp layout

        // ---- zero / surplus ranks: block scan + look-back --------------------
        uint64_t zs_tile;
        const uint64_t zs_excl =
            block_excl_scan<kBlock>(zs_thr, AddOp(), (uint64_t)0, s_u64, &zs_tile);
        if (tid == 0) {
            if (t == 0)
                fs.incl_cnt[t] = zs_tile;
            else
                fs.agg_cnt[t] = zs_tile;
            lb_publish(fs.flags_cnt, t, epoch, t == 0 ? kStIncl : kStAgg);
        }
        if (tid < kWave) {
            uint64_t pre = 0;
            if (t > 0) {
                const uint64_t *agg = fs.agg_cnt, *inc = fs.incl_cnt;
                pre = lookback<uint64_t, uint64_t>(
                    t, fs.flags_cnt, epoch, 0ull, 0ull, [=](int64_t q) { return agg[q]; },
                    [=](int64_t q) { return inc[q]; },
                    [](uint64_t a, uint64_t b) { return a + b; },
                    [](uint64_t a, uint64_t b) { return a + b; }, &s_timeout);
            }
            if (tid == 0) {
                s_bu[2] = pre;
                if (t > 0) {
                    fs.incl_cnt[t] = pre + zs_tile;
                    lb_publish(fs.flags_cnt, t, epoch, kStIncl);
                }
                if (t == fs.tiles - 1) ctl->zeros = (pre + zs_tile) >> 32;
            }
        }
        __syncthreads();
        if (tid == 0 && s_timeout) atomicOr(&ctl->err, (int)kErrTimeout);
        const uint64_t pre = s_bu[2];
        uint64_t zrank = (pre >> 32) + (zs_excl >> 32);
        uint64_t srank_local = zs_excl & 0xffffffffull;  // surplus before this thread (tile)
        const uint64_t s_base = pre & 0xffffffffull;

        // anc identity, zero list, per-item inclusive surplus prefix into LDS
        uint32_t *s_sp = reinterpret_cast<uint32_t *>(s_x);  // kFT entries
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kFI; ++k) {
            const int e = e0 + k;
            const uint64_t i = base + e;
            if (i < M) {
                out.anc[i] = (uint32_t)i;
                if (rep[k] == 0) {
                    if (zrank < M)
                        out.zl[zrank] = (uint32_t)i;
                    else
                        atomicOr(&ctl->err, (int)kErrBounds);
                    ++zrank;
                }
                if (rep[k] >= 2) srank_local += rep[k] - 1;
            }
            s_sp[e] = (uint32_t)srank_local;
        }
        __syncthreads();
        // expand surplus copies: positions [s_base, s_base + S_tile)
        const uint32_t S_tile = (uint32_t)(zs_tile & 0xffffffffull);
        for (uint32_t k = tid; k < S_tile; k += kBlock) {
            int lo = 0, hi = kFT - 1;  // first e with s_sp[e] > k
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_sp[mid] > k)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            if (s_base + k < M)
                out.sl[s_base + k] = (uint32_t)(base + lo);
            else
                atomicOr(&ctl->err, (int)kErrBounds);
        }
    }
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

// residual: integral-part total R (exact integer sum)
template <typename W>
__global__ __launch_bounds__(kBlock) void k_residual_int(FilterCtl *ctl, W w, uint64_t M)
{
    if (!ctl->resample) return;
    __shared__ uint64_t s_u64[kBlock / kWave];
    XResid<W> r{w, 0.0};
    r.bind(ctl);
    uint64_t s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < M;
         i += (uint64_t)gridDim.x * kBlock)
        s += r.ipart(i);
    s = block_reduce<kBlock>(s, AddOp(), s_u64);
    if (threadIdx.x == 0 && s) atomicAdd((unsigned long long *)&ctl->r_int, (unsigned long long)s);
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
    ctl->nn = ctl->n;
    ctl->r_int = 0;
    ctl->delta = 1.0 / (double)ctl->nn;
    if (scheme == kSystematic) ctl->u0 = canonical64(tf64_stream_word(ctl->eng, 0)) * ctl->delta;
}

// ------------------------------------------------------------------------
// sorted uniforms (U01SequenceSorted, u01.hpp:165-202):
//   lmax_n = sum_{k<=n} log(1 - U_k) / (nn - k);  u_n = 1 - exp(lmax_n)
// computed with a fixed reduce-then-scan order (deterministic), then held
// as a tile-wise running max so the sequence is sorted by construction.
// ------------------------------------------------------------------------
__device__ __forceinline__ double sorted_term(const Tf64State &e, uint64_t n, uint64_t nn)
{
    return log(1 - canonical64(tf64_stream_word(e, n))) / (double)(nn - n);
}

__global__ __launch_bounds__(kBlock) void k_sorted_tilesum(FilterCtl *ctl, double *tsum)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kBlock / kWave];
    const uint64_t nn = ctl->nn;
    const uint64_t base = blockIdx.x * (uint64_t)kTile;
    if (base >= nn) return;
    const Tf64State e = ctl->eng;
    double s = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t n = base + threadIdx.x * kItems + k;
        if (n < nn) s += sorted_term(e, n, nn);
    }
    s = block_reduce<kBlock>(s, AddOp(), s_red);
    if (threadIdx.x == 0) tsum[blockIdx.x] = s;
}

// single workgroup: exclusive scan of tile sums (fixed order)
__global__ __launch_bounds__(1024) void k_scan_tiles(FilterCtl *ctl, const double *tsum,
                                                     double *texcl, uint64_t ntiles, int is_max)
{
    if (!ctl->resample) return;
    __shared__ double s_red[1024 / kWave];
    double carry = is_max ? -1.0 : 0.0;
    for (uint64_t b = 0; b < ntiles; b += 1024) {
        const uint64_t i = b + threadIdx.x;
        double v = i < ntiles ? tsum[i] : (is_max ? -1.0 : 0.0);
        double tot;
        double incl = is_max ? block_incl_scan<1024>(v, MaxOp(), s_red, &tot)
                             : block_incl_scan<1024>(v, AddOp(), s_red, &tot);
        double ex = is_max ? (threadIdx.x == 0 ? -1.0 : 0.0) : incl - v;
        if (is_max) {
            double up = lane_up(incl, 1);
            const int lane = threadIdx.x & 63, wid = threadIdx.x / 64;
            // s_red holds per-wave inclusive maxima after the scan
            ex = lane == 0 ? (wid > 0 ? s_red[wid - 1] : -1.0) : up;
            ex = ex < carry ? carry : ex;
        } else {
            ex = carry + ex;
        }
        if (i < ntiles) texcl[i] = ex;
        carry = is_max ? (carry < tot ? tot : carry) : carry + tot;
        __syncthreads();
    }
}

// per tile: lmax = prefix + local inclusive prefix; u = 1 - exp(lmax);
// local inclusive running max of u -> uscan; tile max -> tmaxv
__global__ __launch_bounds__(kBlock) void k_sorted_values(FilterCtl *ctl, const double *texcl,
                                                          double *uscan, double *tmaxv)
{
    if (!ctl->resample) return;
    __shared__ double s_red[kBlock / kWave];
    const uint64_t nn = ctl->nn;
    const uint64_t base = blockIdx.x * (uint64_t)kTile;
    if (base >= nn) return;
    const Tf64State e = ctl->eng;
    double term[kItems];
    double s = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t n = base + threadIdx.x * kItems + k;
        term[k] = n < nn ? sorted_term(e, n, nn) : 0.0;
        s += term[k];
    }
    const double incl = block_incl_scan<kBlock>(s, AddOp(), s_red);
    double lm = texcl[blockIdx.x] + (incl - s);
    double mx = -1.0;
    double uu[kItems];
    for (int k = 0; k < kItems; ++k) {
        lm += term[k];
        const double u = 1 - exp(lm);
        mx = mx < u ? u : mx;
        uu[k] = mx;
    }
    // thread-exclusive running max across the tile
    const double minc = block_incl_scan<kBlock>(mx, MaxOp(), s_red);
    double mex = lane_up(minc, 1);
    {
        const int lane = threadIdx.x & 63, wid = threadIdx.x / 64;
        if (lane == 0) mex = wid > 0 ? s_red[wid - 1] : -1.0;
    }
    for (int k = 0; k < kItems; ++k) {
        const uint64_t n = base + threadIdx.x * kItems + k;
        if (n < nn) uscan[n] = uu[k] < mex ? mex : uu[k];
    }
    if (threadIdx.x == kBlock - 1) tmaxv[blockIdx.x] = minc;
}

// ------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------
// new look-back epoch and a zeroed tile counter for the next k_fold_inv
__global__ void k_fold_begin(FilterCtl *ctl, uint64_t *epoch, unsigned long long *counter)
{
    if (!ctl->resample || threadIdx.x != 0) return;
    *epoch += 1;
    *counter = 0;
}

template <typename XSrc, typename Inv>
static void launch_fold(FilterCtl *ctl, XSrc xs, Inv inv, uint64_t M, FoldScratch &fs,
                        const FoldOut &out, hipStream_t st)
{
    fs.tiles = (M + kFT - 1) / kFT;
    hipLaunchKernelGGL(k_fold_begin, dim3(1), dim3(64), 0, st, ctl, fs.epoch, fs.tile_counter);
    hipLaunchKernelGGL((k_fold_inv<XSrc, Inv>), dim3((unsigned)fs.tiles), dim3(kBlock), 0, st, ctl,
                       xs, inv, M, fs, out);
}

template <typename W>
static void resample_pipeline(FilterCtl *ctl, W w, uint64_t M, int scheme, ResampleBuffers &rb,
                              StepRow *hist, hipStream_t st)
{
    FoldOut out{rb.anc, rb.zl, rb.sl, rb.rep_out, rb.a_out};
    const bool residual = scheme == kResidual || scheme == kResidualStratified ||
                          scheme == kResidualSystematic;
    const unsigned grid = (unsigned)((M + kTile - 1) / kTile);
    if (residual) {
        hipLaunchKernelGGL((k_residual_int<W>), dim3(grid), dim3(kBlock), 0, st, ctl, w, M);
        XResid<W> xr{w, 0.0};
        launch_fold(ctl, xr, InvNone(), M, rb.fs, FoldOut{nullptr, nullptr, nullptr, nullptr, nullptr},
                    st);
        hipLaunchKernelGGL(k_residual_prologue, dim3(1), dim3(64), 0, st, ctl, scheme);
    } else {
        hipLaunchKernelGGL(k_plain_prologue, dim3(1), dim3(64), 0, st, ctl, scheme);
    }
    const bool sorted = scheme == kMultinomial || scheme == kResidual;
    if (sorted) {
        // uniforms: nn <= rb.nuni of them
        const unsigned ut = (unsigned)((rb.nuni + kTile - 1) / kTile);
        hipLaunchKernelGGL(k_sorted_tilesum, dim3(ut), dim3(kBlock), 0, st, ctl, rb.u_tsum);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, ctl, rb.u_tsum, rb.u_texcl,
                           (uint64_t)ut, 0);
        hipLaunchKernelGGL(k_sorted_values, dim3(ut), dim3(kBlock), 0, st, ctl, rb.u_texcl,
                           rb.uscan, rb.u_tmax);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, ctl, rb.u_tmax, rb.u_tmaxex,
                           (uint64_t)ut, 1);
    }
    InvArray ia{rb.uscan, rb.u_tmaxex, 0};
    if (residual) {
        XResidScaled<W> xs{XResid<W>{w, 0.0}, 0.0};
        if (scheme == kResidual)
            launch_fold(ctl, xs, ia, M, rb.fs, out, st);
        else if (scheme == kResidualStratified)
            launch_fold(ctl, xs, InvStratified(), M, rb.fs, out, st);
        else
            launch_fold(ctl, xs, InvSystematic(), M, rb.fs, out, st);
    } else {
        if (scheme == kMultinomial)
            launch_fold(ctl, w, ia, M, rb.fs, out, st);
        else if (scheme == kStratified)
            launch_fold(ctl, w, InvStratified(), M, rb.fs, out, st);
        else
            launch_fold(ctl, w, InvSystematic(), M, rb.fs, out, st);
    }
    if (rb.anc) {
        const unsigned zg = (unsigned)((M + kTile - 1) / kTile) * 4;
        hipLaunchKernelGGL(k_zero_fill, dim3(zg ? zg : 1), dim3(kBlock), 0, st, ctl, rb.zl, rb.sl,
                           rb.anc, M);
    }
    hipLaunchKernelGGL(k_resample_epilogue, dim3(1), dim3(64), 0, st, ctl, hist,
                       (scheme == kSystematic || scheme == kResidualSystematic) ? 1 : 0, M);
}

void launch_resample_from_log_weights(FilterCtl *ctl, const double *lw, uint64_t M, int scheme,
                                      ResampleBuffers &rb, StepRow *hist, hipStream_t st)
{
    resample_pipeline(ctl, XWeight{lw, 0.0, 0.0}, M, scheme, rb, hist, st);
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

}  // namespace vsmcx
