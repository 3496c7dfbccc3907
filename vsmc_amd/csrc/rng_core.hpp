// rng_core.hpp -- counter-based RNG block functions shared by host and device.
//
// Philox4x32-10 follows include/vsmc/rng/philox.hpp:137-265,504-523 (round,
// key bump before rounds 2..10, multipliers 0xD2511F53/0xCD9E8D57, Weyl
// 0x9E3779B9/0xBB67AE85).  Threefry4x64-20 and Threefry4x32-20 follow
// include/vsmc/rng/threefry.hpp:142-233,473-506 (key schedule with the
// 0x1BD11BDA... parity word, injection every 4 rounds, mix pairing by round
// parity).  Engine semantics (counter incremented before each block, words
// returned 0..3, seed(s) -> key {s,0,..}) follow philox.hpp:344-432 and
// threefry.hpp:302-402; the stream helpers below reproduce them from a
// (ctr, key, index) state without a buffer, because buffer == block(ctr)
// whenever index < 4.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VSMCX_HD __host__ __device__ __forceinline__
#else
#define VSMCX_HD inline
#endif

namespace vsmcx {

struct u32x4 { uint32_t v[4]; };
struct u64x4 { uint64_t v[4]; };

VSMCX_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

VSMCX_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    uint32_t s0 = c.v[0], s1 = c.v[1], s2 = c.v[2], s3 = c.v[3];
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32 x 32 -> 64 product gives both halves (v_mad_u64_u32 on gfx950)
        const uint64_t p0 = (uint64_t)0xCD9E8D57u * s2;
        const uint64_t p1 = (uint64_t)0xD2511F53u * s0;
        const uint32_t n0 = (uint32_t)(p0 >> 32) ^ k0 ^ s1;
        const uint32_t n2 = (uint32_t)(p1 >> 32) ^ k1 ^ s3;
        s0 = n0;
        s1 = (uint32_t)p0;
        s2 = n2;
        s3 = (uint32_t)p1;
    }
    u32x4 o;
    o.v[0] = s0; o.v[1] = s1; o.v[2] = s2; o.v[3] = s3;
    return o;
}

VSMCX_HD uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }
VSMCX_HD uint32_t rotl32(uint32_t x, unsigned r) { return (x << r) | (x >> (32 - r)); }

VSMCX_HD u64x4 threefry4x64_20(u64x4 c, const uint64_t key[4])
{
    const unsigned R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                              {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    uint64_t par[5];
    par[4] = 0x1BD11BDAA9FC1A22ull;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint64_t s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = c.v[j] + par[j];
#pragma unroll
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
#pragma unroll
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    u64x4 o;
#pragma unroll
    for (int j = 0; j < 4; ++j) o.v[j] = s[j];
    return o;
}

VSMCX_HD u32x4 threefry4x32_20(u32x4 c, const uint32_t key[4])
{
    const unsigned R[8][2] = {{10, 26}, {11, 21}, {13, 27}, {23, 5},
                              {6, 20}, {17, 11}, {25, 10}, {18, 20}};
    uint32_t par[5];
    par[4] = 0x1BD11BDAu;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        par[j] = key[j];
        par[4] ^= key[j];
    }
    uint32_t s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = c.v[j] + par[j];
#pragma unroll
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
#pragma unroll
            for (int j = 0; j < 4; ++j) s[j] += par[(inc + j) % 5];
            s[3] += inc;
        }
    }
    u32x4 o;
#pragma unroll
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
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t v = (uint64_t)c.v[j] + (carry & 0xffffffffull);
        c.v[j] = (uint32_t)v;
        carry = (carry >> 32) + (v >> 32);
    }
    return c;
}

// Resample engine state: Threefry4x64 with counter, key and buffer index.
// The buffer is block(ctr) whenever index < 4.
struct Tf64State {
    uint64_t ctr[4];
    uint64_t key[4];
    uint32_t index;
    uint32_t pad_;
};

// Word n (0-based) of the stream this engine produces from its current state.
VSMCX_HD uint64_t tf64_stream_word(const Tf64State &e, uint64_t n)
{
    const uint64_t avail = 4u - e.index;
    u64x4 c;
#pragma unroll
    for (int j = 0; j < 4; ++j) c.v[j] = e.ctr[j];
    uint64_t word;
    if (n < avail) {
        word = e.index + n;
    } else {
        const uint64_t m = n - avail;
        c = ctr_add(c, 1 + m / 4);
        word = m % 4;
    }
    const u64x4 b = threefry4x64_20(c, e.key);
    return b.v[word];
}

// Advance the engine by d draws (as d successive operator() calls would).
VSMCX_HD void tf64_advance(Tf64State &e, uint64_t d)
{
    if (d == 0) return;
    const uint64_t avail = 4u - e.index;
    if (d <= avail) {
        e.index += (uint32_t)d;
        return;
    }
    const uint64_t m = d - avail;
    const uint64_t blocks = (m + 3) / 4;
    u64x4 c;
#pragma unroll
    for (int j = 0; j < 4; ++j) c.v[j] = e.ctr[j];
    c = ctr_add(c, blocks);
#pragma unroll
    for (int j = 0; j < 4; ++j) e.ctr[j] = c.v[j];
    e.index = (uint32_t)(m - 4 * (blocks - 1));
}

// std::uniform_real_distribution<double>(0,1) on a 64-bit engine word, as
// libstdc++'s generate_canonical computes it: double(x) / 2^64, a result
// >= 1 replaced by the largest double below 1.
VSMCX_HD double canonical64(uint64_t x)
{
    double r = (double)x / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}

// u01_open_closed_32_53 (include/vsmc/rng/u01.h:139-141)
VSMCX_HD double u01_oc_53(uint32_t u) { return (1.0 + (double)u) * (1.0 / 4294967296.0); }

// u01_open_open_32_24 (include/vsmc/rng/u01.h:116-118): (0.5 + (u >> 9)) 2^-23,
// the midpoint of one of 2^23 equal cells, so its mean is that of the 32-bit
// uniform.  The open-closed 24-bit form (u01.h:112-113) takes the top of each
// 2^-24 cell instead, which lowers E[-2 log u] by about 1e-6: every FP32
// normal from it would be slightly short, a bias shared by all particles.
VSMCX_HD float u01_oo_24(uint32_t u) { return (0.5f + (float)(u >> 9)) * (1.0f / 8388608.0f); }

#if defined(__HIPCC__)
// Box-Muller pair in FP32 on the hardware transcendentals: v_log_f32 is log2,
// v_sin_f32 / v_cos_f32 take the angle in revolutions (sin(2 pi u) = sin_rev(u)).
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float &z0, float &z1)
{
    const float u1 = u01_oo_24(a), u2 = u01_oo_24(b);
    const float r = __builtin_amdgcn_sqrtf(-2.0f * 0.69314718f * __builtin_amdgcn_logf(u1));
    z0 = r * __builtin_amdgcn_cosf(u2);
    z1 = r * __builtin_amdgcn_sinf(u2);
}
#endif

}  // namespace vsmcx
