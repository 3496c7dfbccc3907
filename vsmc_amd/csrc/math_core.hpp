// math_core.hpp -- scalar arithmetic of the hot kernels that is written out
// here rather than taken from the device math library: an FP64 exp for the
// normalised weights and log-sum-exp partials, and the FP32 open-open uniform
// of the Box-Muller normals.  Host and device compile the same IEEE
// operations (fma, rint, ldexp), so a host build gives the device's bits; the
// CPU tests check the host build (tests/cpp/math_host.cpp).
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VSMCX_MHD __host__ __device__ __forceinline__
#else
#define VSMCX_MHD inline
#endif

namespace vsmcx {

// e^d in FP64 for d <= 709 (NaN for NaN, 0 for -inf and below -746).
// Cody and Waite's reduction d = k ln2 + r, |r| <= ln2/2, with ln 2 split in
// two doubles (the low part carries ln 2 - hi), then the Taylor polynomial
// of e^r to degree 13 by Horner's rule (truncation below r^14/14! < 5e-18
// relative) and the scaling 2^k by ldexp, which rounds once into the
// subnormal range.  Within one unit in the last place of e^d (checked on the
// CPU against a correctly rounded exp).  Every constant is an FMA operand, so
// the device keeps them in scalar registers (no vector moves per call).
VSMCX_MHD double exp_lse(double d)
{
    const double k = rint(d * 0x1.71547652b82fep+0);  // d / ln 2
    double r = fma(k, -0x1.62e42fefa39efp-1, d);
    r = fma(k, -0x1.abc9e3b39803fp-56, r);
    double p = 0x1.6124613a86d09p-33;  // 1/13!
    p = fma(p, r, 0x1.1eed8eff8d898p-29);
    p = fma(p, r, 0x1.ae64567f544e4p-26);
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma(p, r, 0x1.71de3a556c734p-19);
    p = fma(p, r, 0x1.a01a01a01a01ap-16);
    p = fma(p, r, 0x1.a01a01a01a01ap-13);
    p = fma(p, r, 0x1.6c16c16c16c17p-10);
    p = fma(p, r, 0x1.1111111111111p-7);
    p = fma(p, r, 0x1.5555555555555p-5);
    p = fma(p, r, 0x1.5555555555555p-3);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const double e = ldexp(p, (int)k);
    return d < -746.0 ? 0.0 : e;
}

// vSMC's u01_open_open_32_24 (include/vsmc/rng/u01.h:116-118),
// ((u >> 9) + 0.5) 2^-23, built from the bits: 0x3F800000 | (u >> 9) is
// 1 + (u >> 9) 2^-23 in [1, 2), and subtracting 1 - 2^-24 is exact (Sterbenz:
// both operands lie within a factor 2) and leaves (2 (u >> 9) + 1) 2^-24,
// which has 24 significant bits.  On the device the first step is one
// funnel shift of {0x7F, u} right by 9.
VSMCX_MHD float u01_oo_24_bits(uint32_t u)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t b = __builtin_amdgcn_alignbit(0x7Fu, u, 9u);
#else
    const uint32_t b = 0x3F800000u | (u >> 9);
#endif
    float f;
    memcpy(&f, &b, sizeof f);
    return f - 0x1.fffffep-1f;
}

}  // namespace vsmcx
