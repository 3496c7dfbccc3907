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

// One fold step a' = fl(a + x) for a inside binade b, as a transducer.
VSMCX_HD Tr tr_element(double x, int b)
{
    const uint64_t bits = dbits(x);
    const int ex = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    const uint64_t mx = ex ? (frac | (1ull << 52)) : frac;
    const int exe = ex ? ex : 1;
    const int s = (b + 1023) - exe;  // bits of mx below the grid
    uint64_t q;
    int kind;  // 0 down, 1 up, 2 tie
    if (s <= 0) {
        q = (-s >= 7) ? kSat : (mx << (-s));
        kind = 0;
    } else if (s >= 55) {
        q = 0;
        kind = 0;
    } else {
        q = mx >> s;
        const uint64_t rem = mx & ((1ull << s) - 1);
        const uint64_t half = 1ull << (s - 1);
        kind = rem > half ? 1 : (rem == half ? 2 : 0);
    }
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

// Binade that every value within the error bound of S falls in, or kNoBinade.
VSMCX_HD int certain_binade(double S, double eps_rel, double eps_abs)
{
    const double err = S * eps_rel + eps_abs;
    double lo = S - err;
    if (lo < 0) lo = 0;
    const double hi = S + err;
    const int bl = binade_of(lo), bh = binade_of(hi);
    return bl == bh ? bl : kNoBinade;
}

// Error bound of an approximate prefix sum of n non-negative terms against the
// left fold over the same terms: both deviate from the exact real prefix by at
// most ~n ulps plus a few ulps per term for a recomputed term; the factor 2^6
// is headroom.  Too large a bound only costs more specials.
VSMCX_HD void fold_eps(uint64_t n, double *eps_rel, double *eps_abs)
{
    const double nn = (double)n + 65536.0;
    *eps_rel = nn * 0x1p-47;
    *eps_abs = nn * 0x1p-1068;
}

}  // namespace vsmcx
