// vsmcx_device.hpp -- counter-based RNGs for HIP kernels, user moves included.
//
// The block functions are the ones the library's own kernels use:
// Philox4x32-10 (include/vsmc/rng/philox.hpp:137-265,504-523: multipliers
// 0xD2511F53 / 0xCD9E8D57, Weyl 0x9E3779B9 / 0xBB67AE85, key bumped before
// rounds 2..10), Threefry4x64-20 and Threefry4x32-20
// (include/vsmc/rng/threefry.hpp:142-233,473-506: key schedule with the
// 0x1BD11BDA... parity word, injection every 4 rounds).  They match the
// Random123 known-answer vectors (tests/test_oracle.py).
//
// Above them, the OpenCL-C RNG layer of vSMC, for per-particle moves written
// as device code (the convention of example/pf/pf_cl.cl):
//   * urng (include/vsmc/rng/urng.h:179-226): <rng>_rng_t {key, ctr, rnd,
//     remain}; <rng>_init(rng, seed) sets key {seed, 0, ...} and a zero
//     counter; <rng>_rand(rng) bumps word 0 of the counter (no carry) when
//     the block is used up and returns the block's words last to first;
//     cburng4x32 / cburng4x64 are Threefry unless VSMC_USE_PHILOX_CBURNG is
//     defined non-zero (urng.h:282-332);
//   * u01 (include/vsmc/rng/u01.h:103-211): the twelve integer-to-uniform
//     conversions;
//   * normal01 (include/vsmc/rng/normal01.h:118-139): Box-Muller with vSMC's
//     draw order -- init draws (u1, u2) from (0, 1]; rand returns
//     sqrt(-2 log u1) cos(2 pi u2) the first time, then draws a new pair and
//     returns sqrt(-2 log u1) sin(2 pi u2), and so on.
// Header only; host code may include it too (the functions are host and
// device).
#ifndef VSMCX_DEVICE_HPP
#define VSMCX_DEVICE_HPP

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VSMCX_HD __host__ __device__ __forceinline__
#else
#define VSMCX_HD inline
#endif
#if defined(__clang__)
#define VSMCX_UNROLL _Pragma("unroll")
#else
#define VSMCX_UNROLL
#endif

namespace vsmcx {

struct u32x2 { uint32_t v[2]; };
struct u32x4 { uint32_t v[4]; };
struct u64x4 { uint64_t v[4]; };

// ---- block functions ----------------------------------------------------------------
VSMCX_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// a ^ b ^ c: one v_bitop3_b32 on the device (gfx950's three-input bit
// operation, truth table 0x96), two XORs on the host
VSMCX_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

VSMCX_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    uint32_t s0 = c.v[0], s1 = c.v[1], s2 = c.v[2], s3 = c.v[3];
    VSMCX_UNROLL
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32 x 32 -> 64 product gives both halves (v_mad_u64_u32 on gfx950)
        const uint64_t p0 = (uint64_t)0xCD9E8D57u * s2;
        const uint64_t p1 = (uint64_t)0xD2511F53u * s0;
        const uint32_t n0 = xor3((uint32_t)(p0 >> 32), s1, k0);
        const uint32_t n2 = xor3((uint32_t)(p1 >> 32), s3, k1);
        s0 = n0;
        s1 = (uint32_t)p0;
        s2 = n2;
        s3 = (uint32_t)p1;
    }
    u32x4 o;
    o.v[0] = s0; o.v[1] = s1; o.v[2] = s2; o.v[3] = s3;
    return o;
}

// Rotations by a constant in 1..63.  On the device the two 32-bit halves
// come from one funnel shift each (v_alignbit_b32), with the halves swapped
// for r >= 32.
VSMCX_HD uint64_t rotl64(uint64_t x, unsigned r)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t a = r < 32 ? hi : lo, b = r < 32 ? lo : hi;  // {a, b} = x rotated by 0 or 32
    const unsigned s = r & 31u;
    if (s == 0) return ((uint64_t)a << 32) | b;
    const uint32_t nh = __builtin_amdgcn_alignbit(a, b, 32u - s);
    const uint32_t nl = __builtin_amdgcn_alignbit(b, a, 32u - s);
    return ((uint64_t)nh << 32) | nl;
#else
    return (x << r) | (x >> (64 - r));
#endif
}
VSMCX_HD uint32_t rotl32(uint32_t x, unsigned r) { return (x << r) | (x >> (32 - r)); }

VSMCX_HD u64x4 threefry4x64_20(u64x4 c, const uint64_t key[4])
{
    const unsigned R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                              {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    uint64_t par[5];
    par[4] = 0x1BD11BDAA9FC1A22ull;
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint64_t s[4];
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) s[j] = c.v[j] + par[j];
    VSMCX_UNROLL
    for (unsigned n = 1; n <= 20; ++n) {
        const unsigned r = (n - 1) % 8;
        const unsigned i0 = (n % 2) ? 1 : 3, i2 = (n % 2) ? 3 : 1;
        s[0] += s[i0];
        s[i0] = rotl64(s[i0], R[r][0]);
        s[i0] ^= s[0];
        s[2] += s[i2];
        s[i2] = rotl64(s[i2], R[r][1]);
        s[i2] ^= s[2];
        if (n % 4 == 0) {
            const unsigned inc = n / 4;
            VSMCX_UNROLL
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    u64x4 o;
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) o.v[j] = s[j];
    return o;
}

VSMCX_HD u32x4 threefry4x32_20(u32x4 c, const uint32_t key[4])
{
    const unsigned R[8][2] = {{10, 26}, {11, 21}, {13, 27}, {23, 5},
                              {6, 20}, {17, 11}, {25, 10}, {18, 20}};
    uint32_t par[5];
    par[4] = 0x1BD11BDAu;
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint32_t s[4];
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) s[j] = c.v[j] + par[j];
    VSMCX_UNROLL
    for (unsigned n = 1; n <= 20; ++n) {
        const unsigned r = (n - 1) % 8;
        const unsigned i0 = (n % 2) ? 1 : 3, i2 = (n % 2) ? 3 : 1;
        s[0] += s[i0];
        s[i0] = rotl32(s[i0], R[r][0]);
        s[i0] ^= s[0];
        s[2] += s[i2];
        s[i2] = rotl32(s[i2], R[r][1]);
        s[i2] ^= s[2];
        if (n % 4 == 0) {
            const unsigned inc = n / 4;
            VSMCX_UNROLL
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    u32x4 o;
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) o.v[j] = s[j];
    return o;
}

// ctr + n for a 4-word counter of 64-bit words (utility/counter.hpp:128-152)
VSMCX_HD u64x4 ctr_add(u64x4 c, uint64_t n)
{
    const uint64_t old = c.v[0];
    c.v[0] += n;
    if (c.v[0] < old) {
        if (++c.v[1] == 0)
            if (++c.v[2] == 0) ++c.v[3];
    }
    return c;
}

// ctr + n for a 4-word counter of 32-bit words
VSMCX_HD u32x4 ctr_add(u32x4 c, uint64_t n)
{
    uint64_t carry = n;
    VSMCX_UNROLL
    for (int j = 0; j < 4; ++j) {
        const uint64_t v = (uint64_t)c.v[j] + (carry & 0xffffffffull);
        c.v[j] = (uint32_t)v;
        carry = (carry >> 32) + (v >> 32);
    }
    return c;
}

// ---- urng: one engine per work item --------------------------------------------------
struct philox4x32_rng_t {
    u32x2 key;
    u32x4 ctr;
    u32x4 rnd;
    unsigned char remain;
};
struct threefry4x32_rng_t {
    u32x4 key;
    u32x4 ctr;
    u32x4 rnd;
    unsigned char remain;
};
struct threefry4x64_rng_t {
    u64x4 key;
    u64x4 ctr;
    u64x4 rnd;
    unsigned char remain;
};

VSMCX_HD void philox4x32_init(philox4x32_rng_t *rng, uint32_t seed)
{
    rng->key.v[0] = seed;
    rng->key.v[1] = 0;
    for (int j = 0; j < 4; ++j) rng->ctr.v[j] = rng->rnd.v[j] = 0;
    rng->remain = 0;
}
VSMCX_HD void threefry4x32_init(threefry4x32_rng_t *rng, uint32_t seed)
{
    for (int j = 0; j < 4; ++j) rng->key.v[j] = rng->ctr.v[j] = rng->rnd.v[j] = 0;
    rng->key.v[0] = seed;
    rng->remain = 0;
}
VSMCX_HD void threefry4x64_init(threefry4x64_rng_t *rng, uint64_t seed)
{
    for (int j = 0; j < 4; ++j) rng->key.v[j] = rng->ctr.v[j] = rng->rnd.v[j] = 0;
    rng->key.v[0] = seed;
    rng->remain = 0;
}

VSMCX_HD uint32_t philox4x32_rand(philox4x32_rng_t *rng)
{
    if (rng->remain > 0) return rng->rnd.v[--rng->remain];
    ++rng->ctr.v[0];
    rng->rnd = philox4x32_10(rng->ctr, rng->key.v[0], rng->key.v[1]);
    rng->remain = 3;
    return rng->rnd.v[3];
}
VSMCX_HD uint32_t threefry4x32_rand(threefry4x32_rng_t *rng)
{
    if (rng->remain > 0) return rng->rnd.v[--rng->remain];
    ++rng->ctr.v[0];
    rng->rnd = threefry4x32_20(rng->ctr, rng->key.v);
    rng->remain = 3;
    return rng->rnd.v[3];
}
VSMCX_HD uint64_t threefry4x64_rand(threefry4x64_rng_t *rng)
{
    if (rng->remain > 0) return rng->rnd.v[--rng->remain];
    ++rng->ctr.v[0];
    rng->rnd = threefry4x64_20(rng->ctr, rng->key.v);
    rng->remain = 3;
    return rng->rnd.v[3];
}

#if defined(VSMC_USE_PHILOX_CBURNG) && VSMC_USE_PHILOX_CBURNG
typedef philox4x32_rng_t cburng4x32_rng_t;
VSMCX_HD void cburng4x32_init(cburng4x32_rng_t *r, uint32_t s) { philox4x32_init(r, s); }
VSMCX_HD uint32_t cburng4x32_rand(cburng4x32_rng_t *r) { return philox4x32_rand(r); }
#else
typedef threefry4x32_rng_t cburng4x32_rng_t;
VSMCX_HD void cburng4x32_init(cburng4x32_rng_t *r, uint32_t s) { threefry4x32_init(r, s); }
VSMCX_HD uint32_t cburng4x32_rand(cburng4x32_rng_t *r) { return threefry4x32_rand(r); }
#endif
typedef threefry4x64_rng_t cburng4x64_rng_t;
VSMCX_HD void cburng4x64_init(cburng4x64_rng_t *r, uint64_t s) { threefry4x64_init(r, s); }
VSMCX_HD uint64_t cburng4x64_rand(cburng4x64_rng_t *r) { return threefry4x64_rand(r); }

// ---- u01: integers to uniforms -------------------------------------------------------
// 32 bits to float (24 bits of resolution; closed-closed keeps 25 and rounds)
VSMCX_HD float u01_closed_closed_32_24(uint32_t u)
{
    return (float)((u & 0x7fffffc0u) + (u & 0x40u)) * (1.0f / 2147483648.0f);
}
VSMCX_HD float u01_closed_open_32_24(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }
VSMCX_HD float u01_open_closed_32_24(uint32_t u) { return (1.0f + (float)(u >> 8)) * (1.0f / 16777216.0f); }
VSMCX_HD float u01_open_open_32_24(uint32_t u) { return (0.5f + (float)(u >> 9)) * (1.0f / 8388608.0f); }
// 32 bits to double
VSMCX_HD double u01_closed_closed_32_53(uint32_t u)
{
    return ((double)(u & 1u) + (double)u) * (1.0 / 4294967296.0);
}
VSMCX_HD double u01_closed_open_32_53(uint32_t u) { return (double)u * (1.0 / 4294967296.0); }
VSMCX_HD double u01_open_closed_32_53(uint32_t u) { return (1.0 + (double)u) * (1.0 / 4294967296.0); }
VSMCX_HD double u01_open_open_32_53(uint32_t u) { return (0.5 + (double)u) * (1.0 / 4294967296.0); }
// 64 bits to double
VSMCX_HD double u01_closed_closed_64_53(uint64_t u)
{
    return (double)((u & 0x7ffffffffffffe00ull) + (u & 0x200ull)) * (1.0 / 9223372036854775808.0);
}
VSMCX_HD double u01_closed_open_64_53(uint64_t u) { return (double)(u >> 11) * (1.0 / 9007199254740992.0); }
VSMCX_HD double u01_open_closed_64_53(uint64_t u)
{
    return (1.0 + (double)(u >> 11)) * (1.0 / 9007199254740992.0);
}
VSMCX_HD double u01_open_open_64_53(uint64_t u)
{
    return (0.5 + (double)(u >> 12)) * (1.0 / 4503599627370496.0);
}

// ---- normal01: Box-Muller in vSMC's draw order ---------------------------------------
struct normal01_4x32_24 { float u1, u2; unsigned char saved; };
struct normal01_4x32_53 { double u1, u2; unsigned char saved; };
struct normal01_4x64_53 { double u1, u2; unsigned char saved; };

VSMCX_HD void normal01_4x32_24_init(normal01_4x32_24 *n, cburng4x32_rng_t *rng)
{
    n->u1 = u01_open_closed_32_24(cburng4x32_rand(rng));
    n->u2 = u01_open_closed_32_24(cburng4x32_rand(rng));
    n->saved = 1;
}
VSMCX_HD float normal01_4x32_24_rand(normal01_4x32_24 *n, cburng4x32_rng_t *rng)
{
    const float c_2pi = 6.2831853071795865f;
    if (n->saved) {
        n->saved = 0;
        return sqrtf(-2 * logf(n->u1)) * cosf(c_2pi * n->u2);
    }
    normal01_4x32_24_init(n, rng);
    return sqrtf(-2 * logf(n->u1)) * sinf(c_2pi * n->u2);
}
VSMCX_HD void normal01_4x32_53_init(normal01_4x32_53 *n, cburng4x32_rng_t *rng)
{
    n->u1 = u01_open_closed_32_53(cburng4x32_rand(rng));
    n->u2 = u01_open_closed_32_53(cburng4x32_rand(rng));
    n->saved = 1;
}
VSMCX_HD double normal01_4x32_53_rand(normal01_4x32_53 *n, cburng4x32_rng_t *rng)
{
    const double c_2pi = 6.2831853071795865;
    if (n->saved) {
        n->saved = 0;
        return sqrt(-2 * log(n->u1)) * cos(c_2pi * n->u2);
    }
    normal01_4x32_53_init(n, rng);
    return sqrt(-2 * log(n->u1)) * sin(c_2pi * n->u2);
}
VSMCX_HD void normal01_4x64_53_init(normal01_4x64_53 *n, cburng4x64_rng_t *rng)
{
    n->u1 = u01_open_closed_64_53(cburng4x64_rand(rng));
    n->u2 = u01_open_closed_64_53(cburng4x64_rand(rng));
    n->saved = 1;
}
VSMCX_HD double normal01_4x64_53_rand(normal01_4x64_53 *n, cburng4x64_rng_t *rng)
{
    const double c_2pi = 6.2831853071795865;
    if (n->saved) {
        n->saved = 0;
        return sqrt(-2 * log(n->u1)) * cos(c_2pi * n->u2);
    }
    normal01_4x64_53_init(n, rng);
    return sqrt(-2 * log(n->u1)) * sin(c_2pi * n->u2);
}

}  // namespace vsmcx

#endif  // VSMCX_DEVICE_HPP
