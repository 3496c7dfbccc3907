// fold_core.hpp -- exact parallel evaluation of a sequential FP64 left fold.
//
// vSMC's inversion (include/vsmc/resample/common.hpp:76-83) and asum
// (include/vsmc/math/cblas.hpp:53-63) are left folds a_i = fl(a_{i-1} + x_i)
// over non-negative x.  A parallel tree sum rounds differently, and at 2^24
// particles that moves tens of ancestors per resample.  This header turns the
// fold into an associative problem:
//
//  * While the running value stays inside one binade [2^b, 2^(b+1)) every
//    partial sum lies on the grid g = 2^(b-52).  Writing a = A*g and
//    x = g*(q + f), round-to-nearest-even gives a' = (A + q + r)*g with r = 0
//    (f < 1/2), 1 (f > 1/2), or, on a tie, r = (A + q) odd.  So one element
//    is a "transducer" over the parity of A: increment d_p and new parity
//    P(p) for p in {0, 1}.  Transducers compose associatively (function
//    composition), so a segment of elements in one binade reduces to
//    (d0, d1, P0, P1) exactly, with integer arithmetic.
//  * Elements where the running value may cross a binade boundary are
//    "specials": evaluated with one real FP64 add, in order.  An approximate
//    prefix sum with a proven error bound classifies every element; a wrong
//    guess can only make an element special, never make a regular element
//    wrong.
#pragma once
#include <stdint.h>

#include "rng_core.hpp"

namespace vsmcx {

constexpr int kNoBinade = 0x40000000;
constexpr uint64_t kSat = 1ull << 60;
constexpr uint64_t kTwo53 = 1ull << 53;

VSMCX_HD uint64_t dbits(double v)
{
    union { double d; uint64_t u; } c;
    c.d = v;
    return c.u;
}
VSMCX_HD double bitsd(uint64_t u)
{
    union { double d; uint64_t u; } c;
    c.u = u;
    return c.d;
}

// binade index b of v >= 0: v in [2^b, 2^(b+1)), grid 2^(b-52); zero and
// subnormals share b = -1022 (grid 2^-1074).
VSMCX_HD int binade_of(double v)
{
    int e = (int)((dbits(v) >> 52) & 0x7ff);
    return (e < 1 ? 1 : e) - 1023;
}

// v / 2^(binade_of(v) - 52), exact.
VSMCX_HD uint64_t units_of(double v)
{
    const uint64_t b = dbits(v);
    const uint64_t e = (b >> 52) & 0x7ff;
    const uint64_t f = b & ((1ull << 52) - 1);
    return e ? (f | (1ull << 52)) : f;
}

// The double A * 2^(b-52) for A < 2^53 (exact).
VSMCX_HD double value_of(int b, uint64_t A)
{
    return bitsd(((uint64_t)(b + 1022) << 52) + A);
}

// Parity transducer: from parity p the segment adds d[p] grid units and
// leaves parity P(p) = (pp >> p) & 1.
struct Tr {
    uint64_t d0, d1;
    uint32_t pp;
    int32_t b;  // binade the segment lives in, kNoBinade if empty
};

VSMCX_HD Tr tr_identity()
{
    Tr t;
    t.d0 = 0;
    t.d1 = 0;
    t.pp = 2u;
    t.b = kNoBinade;
    return t;
}

VSMCX_HD uint64_t sat_add(uint64_t a, uint64_t b)
{
    const uint64_t s = a + b;
    return s > kSat ? kSat : s;
}

// branch-free (p ? b : a) for p in {0, 1}; keeps the pair in registers
VSMCX_HD uint64_t pick(uint64_t a, uint64_t b, uint64_t p)
{
    return a ^ ((a ^ b) & (0ull - p));
}

// f then g.  Binade tags must agree when both are non-empty; a mismatch is
// reported by tagging the result with kNoBinade - 1.
VSMCX_HD Tr tr_compose(const Tr &f, const Tr &g)
{
    Tr r;
    const uint32_t f0 = f.pp & 1u, f1 = (f.pp >> 1) & 1u;
    r.d0 = sat_add(f.d0, pick(g.d0, g.d1, f0));
    r.d1 = sat_add(f.d1, pick(g.d0, g.d1, f1));
    const uint32_t p0 = (g.pp >> f0) & 1u, p1 = (g.pp >> f1) & 1u;
    r.pp = p0 | (p1 << 1);
    if (f.b == kNoBinade)
        r.b = g.b;
    else if (g.b == kNoBinade || g.b == f.b)
        r.b = f.b;
    else
        r.b = kNoBinade - 1;
    return r;
}

// x in grid units of binade b: integral part q (saturated at kSat) and how
// the fraction rounds: 0 down, 1 up, 2 tie.
VSMCX_HD uint64_t grid_units(double x, int b, int *kind)
{
    const uint64_t bits = dbits(x);
    const int ex = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    const uint64_t mx = ex ? (frac | (1ull << 52)) : frac;
    const int exe = ex ? ex : 1;
    const int s = (b + 1023) - exe;  // bits of mx below the grid
    if (s <= 0) {
        *kind = 0;
        return (-s >= 7) ? kSat : (mx << (-s));
    }
    if (s >= 55) {
        *kind = 0;
        return 0;
    }
    const uint64_t rem = mx & ((1ull << s) - 1);
    const uint64_t half = 1ull << (s - 1);
    *kind = rem > half ? 1 : (rem == half ? 2 : 0);
    return mx >> s;
}

// One fold step a' = fl(a + x) for a inside binade b, as a transducer.
VSMCX_HD Tr tr_element(double x, int b)
{
    int kind;
    const uint64_t q = grid_units(x, b, &kind);
    Tr t;
    t.b = b;
    if (kind == 0) {
        t.d0 = q;
        t.d1 = q;
    } else if (kind == 1) {
        t.d0 = q + 1;
        t.d1 = q + 1;
    } else {
        t.d0 = q + (q & 1);
        t.d1 = q + 1 - (q & 1);
    }
    if (t.d0 > kSat) t.d0 = kSat;
    if (t.d1 > kSat) t.d1 = kSat;
    const uint32_t p0 = (uint32_t)(t.d0 & 1), p1 = (uint32_t)((1 + t.d1) & 1);
    t.pp = p0 | (p1 << 1);
    return t;
}

// A run of elements in binade b none of which is a tie adds the same number
// of grid units d whatever the parity: the transducer (d, d, parity flip by d).
VSMCX_HD Tr tr_units(uint64_t d, int b)
{
    Tr t;
    t.d0 = d > kSat ? kSat : d;
    t.d1 = t.d0;
    t.pp = (uint32_t)(t.d0 & 1) | ((uint32_t)((1 + t.d0) & 1) << 1);
    t.b = b;
    return t;
}

#if defined(__HIPCC__)
// field-wise cross-lane moves (see device_util.hpp)
__device__ __forceinline__ Tr lane_up(const Tr &v, unsigned d)
{
    Tr r;
    r.d0 = (uint64_t)__shfl_up((unsigned long long)v.d0, d, 64);
    r.d1 = (uint64_t)__shfl_up((unsigned long long)v.d1, d, 64);
    r.pp = (uint32_t)__shfl_up((int)v.pp, d, 64);
    r.b = __shfl_up(v.b, d, 64);
    return r;
}
__device__ __forceinline__ Tr lane_get(const Tr &v, int l)
{
    Tr r;
    r.d0 = (uint64_t)__shfl((unsigned long long)v.d0, l, 64);
    r.d1 = (uint64_t)__shfl((unsigned long long)v.d1, l, 64);
    r.pp = (uint32_t)__shfl((int)v.pp, l, 64);
    r.b = __shfl(v.b, l, 64);
    return r;
}
#endif

// Apply a transducer to an exact running value v.  Returns false if v is not
// in the segment's binade or the result would leave it (never happens when
// the classification bound holds).
VSMCX_HD bool tr_apply(const Tr &t, double v, double *out)
{
    if (t.b == kNoBinade) {
        *out = v;
        return true;
    }
    if (t.b == kNoBinade - 1 || binade_of(v) != t.b) return false;
    const uint64_t A = units_of(v);
    const uint64_t A2 = A + pick(t.d0, t.d1, A & 1);
    if (A2 >= kTwo53) return false;
    *out = value_of(t.b, A2);
    return true;
}

// Equal values, NaN matching NaN (a fold that met a NaN stays NaN).
VSMCX_HD bool same_value(double a, double b) { return a == b || (a != a && b != b); }

// Lower end of the error interval around an approximate prefix S.
VSMCX_HD double err_lo(double S, double eps_rel, double eps_abs)
{
    const double lo = S - (S * eps_rel + eps_abs);
    return lo < 0 ? 0 : lo;
}

// Binade that every value within the error bound of S falls in, or kNoBinade.
VSMCX_HD int certain_binade(double S, double eps_rel, double eps_abs)
{
    const double hi = S + (S * eps_rel + eps_abs);
    const int bl = binade_of(err_lo(S, eps_rel, eps_abs)), bh = binade_of(hi);
    return bl == bh ? bl : kNoBinade;
}

// ---- segmented form of the whole fold -------------------------------------------
// Every element is one of
//   regular:     a_{i-1} and a_i lie for certain in one binade b -> its transducer;
//   negligible:  x_i = 0, or x_i is below half the smallest grid a_{i-1}
//                can have, so a_i = a_{i-1} exactly               -> identity;
//   special:     anything else; evaluated later with one real FP64 add.
// A Seg is the fold of a run of elements: the number of specials in it and the
// transducer of the elements after its last special.  Regular elements between
// two specials share one binade (a binade change forces a special), so the
// composition never mixes binades when the classification bound holds.
struct Seg {
    Tr tr;
    uint64_t cnt;
};

VSMCX_HD Seg seg_identity()
{
    Seg s;
    s.tr = tr_identity();
    s.cnt = 0;
    return s;
}

// a then b
VSMCX_HD Seg seg_op(const Seg &a, const Seg &b)
{
    Seg r;
    const Tr c = tr_compose(a.tr, b.tr);
    const uint64_t p = b.cnt != 0;
    r.tr.d0 = pick(c.d0, b.tr.d0, p);
    r.tr.d1 = pick(c.d1, b.tr.d1, p);
    r.tr.pp = p ? b.tr.pp : c.pp;
    r.tr.b = p ? b.tr.b : c.b;
    r.cnt = a.cnt + b.cnt;
    return r;
}

// Classify element x with approximate prefixes sb (before) and sa (after).
// x = 0 never changes the running value (also past the end of the data, and
// after a NaN or infinity); once the approximate prefix is not finite every
// further non-zero element is special, which gives IEEE propagation.
VSMCX_HD Seg elem_seg(double x, double sb, double sa, double eps_rel, double eps_abs)
{
    Seg s = seg_identity();
    if (x == 0.0) return s;
    if (!(binade_of(sa) < 1024)) {  // sa = sb + x is NaN or infinite
        s.cnt = 1;
        return s;
    }
    const int bp = certain_binade(sb, eps_rel, eps_abs);
    const int bc = certain_binade(sa, eps_rel, eps_abs);
    if (bp == bc && bp != kNoBinade) {
        s.tr = tr_element(x, bc);
    } else {
        // a_{i-1} >= 2^blo, so its grid is at least 2^(blo-52); x < 2^(blo-53)
        // (implied by binade_of(x) <= blo - 54) rounds away
        const int blo = binade_of(err_lo(sb, eps_rel, eps_abs));
        s.cnt = binade_of(x) <= blo - 54 ? 0 : 1;
    }
    return s;
}

// ---- one thread's run of consecutive elements -----------------------------------
// x(k), k < n, are the elements; s is the approximate prefix before the first
// (the per-element prefixes are s + x(0), then + x(1), ...).
//
// Fast path: when the error intervals around the first and the last prefix lie
// in one binade b, every element is regular in b (the prefix and both ends of
// its interval are monotone), and without ties the run adds the same number of
// grid units whatever the parity.  fast_b receives b, or kNoBinade for the
// element-by-element path.
//
// The two paths are also callable apart (run_seg_fast, run_seg_elems), so that
// device code can run the fast path over values held in registers (its loops
// unroll) and the rare element-by-element path over values in LDS.
//
// The run's units d come from its own fold steps started at 2^b (b above the
// subnormal binade): without ties each step adds q_k + [fraction > 1/2] grid
// units whatever the running value in the binade, the run's increments total
// less than 2^b (the exact run lies in [2^b, 2^(b+1))), so from 2^b the
// steps stay in the binade and d is the significand field of the end value.
// A step is a tie when its rounding error, x_k - (S_k - S_{k-1}) (exact: the
// fast two-sum with S_{k-1} >= 2^b > x_k), is half a grid unit.
template <typename X>
VSMCX_HD bool run_seg_fast(X x, int n, double s, double er, double ea, int *fast_b, Seg *r)
{
    double se = s;
    for (int k = 0; k < n; ++k) se += x(k);
    const int b0 = certain_binade(s, er, ea);
    if (!(b0 < 1024 && b0 == certain_binade(se, er, ea))) return false;
    uint64_t d = 0;
    int tie = 0;
    if (b0 > -1022) {
        const double half = ldexp(1.0, b0 - 53);
        double S = value_of(b0, 1ull << 52);  // 2^b0
        for (int k = 0; k < n; ++k) {
            const double xk = x(k);
            const double Sn = S + xk;
            tie |= fabs(xk - (Sn - S)) == half;
            S = Sn;
        }
        d = dbits(S) & ((1ull << 52) - 1);
    } else {
        for (int k = 0; k < n; ++k) {
            int kind;
            d += grid_units(x(k), b0, &kind) + (kind == 1);
            tie |= kind == 2;
        }
    }
    if (tie) return false;
    *fast_b = b0;
    r->tr = tr_units(d, b0);
    r->cnt = 0;
    return true;
}

template <typename X>
VSMCX_HD Seg run_seg_elems(X x, int n, double s, double er, double ea)
{
    Seg g = seg_identity();
    // the element-by-element path is rare: kept rolled, so the common path
    // stays small in the instruction cache
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll 1
#endif
    for (int k = 0; k < n; ++k) {
        const double xk = x(k);
        const double sb = s;
        s = sb + xk;
        g = seg_op(g, elem_seg(xk, sb, s, er, ea));
    }
    return g;
}

template <typename X>
VSMCX_HD Seg run_seg(X x, int n, double s, double er, double ea, int *fast_b)
{
    Seg r;
    if (run_seg_fast(x, n, s, er, ea, fast_b, &r)) return r;
    *fast_b = kNoBinade;
    return run_seg_elems(x, n, s, er, ea);
}

// The exact value before a run: pre is the summary of everything before it,
// vb(c) the exact value right after special c (fold order).  From there the
// run's own values are vSMC's fold steps a = fl(a + x), exact by construction;
// a run that ends where the next run's start says it should confirms that
// start (and a chain of such checks confirms every value).
struct RunStart {
    uint64_t vcnt;  // specials before the run
    double vbase;   // exact value right after the last of them (0 if none)
    double a_prev;  // exact value before the run's first element
    bool ok;
};

template <typename VB>
VSMCX_HD RunStart run_start(const Seg &pre, VB vb)
{
    RunStart r;
    r.ok = true;
    r.vcnt = pre.cnt;
    r.vbase = r.vcnt ? vb(r.vcnt - 1) : 0.0;
    if (!tr_apply(pre.tr, r.vbase, &r.a_prev)) {
        r.ok = false;
        r.a_prev = r.vbase;
    }
    return r;
}

// One special element in fold order: the run transducer since the previous
// special (or since the start, value 0) and x; the chain pass overwrites x
// with the exact fold value right after the element.
struct SpecRec {
    Tr run;
    double x;
};

#if defined(__HIPCC__)
__device__ __forceinline__ Seg lane_up(const Seg &v, unsigned d)
{
    Seg r;
    r.tr = lane_up(v.tr, d);
    r.cnt = (uint64_t)__shfl_up((unsigned long long)v.cnt, d, 64);
    return r;
}
__device__ __forceinline__ Seg lane_get(const Seg &v, int l)
{
    Seg r;
    r.tr = lane_get(v.tr, l);
    r.cnt = (uint64_t)__shfl((unsigned long long)v.cnt, l, 64);
    return r;
}
#endif

// Error bound of an approximate prefix sum of n non-negative terms against the
// left fold over the same terms.  Any summation order of n non-negative terms
// is within gamma_{n-1} * sum of the exact sum (Higham, Accuracy and Stability
// of Numerical Algorithms, 2nd ed., eq. 4.4), so the two differ by at most
// 2 * n * 2^-53 * sum; the bound below is twice that, and an absolute term
// covers subnormal partial sums.  The tile starts may also come from terms
// that differ from the fold's own x_i by at most 1500 u relative each (the
// filter's tile sums from the move's log-sum-exp partials); the 512 extra
// units below (2048 u) cover that on their own.  Too large a bound only costs
// more specials; too small breaks exactness, which the device reports
// (tr_apply fails) instead of returning wrong values.
VSMCX_HD void fold_eps(uint64_t n, double *eps_rel, double *eps_abs)
{
    const double nn = (double)n + 576.0;
    *eps_rel = nn * 0x1p-51;
    *eps_abs = nn * 0x1p-1070;
}

}  // namespace vsmcx
