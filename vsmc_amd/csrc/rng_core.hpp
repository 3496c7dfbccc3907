// rng_core.hpp -- the library's engine helpers above the block functions
// (Philox4x32-10, Threefry4x64-20, Threefry4x32-20, counters and u01
// conversions, in include/vsmcx_device.hpp, which user kernels include too).
//
// The resampler's Threefry4x64 engine follows vSMC's engine semantics
// (counter incremented before each block, words returned 0..3, seed(s) ->
// key {s,0,..}; include/vsmc/rng/threefry.hpp:302-402): the stream helpers
// below reproduce them from a (ctr, key, index) state without a buffer,
// because buffer == block(ctr) whenever index < 4.
#pragma once
#include <stdint.h>

// the block functions, counters and u01 conversions (also the public device
// header for user kernels)
#include "../../include/vsmcx_device.hpp"
#include "math_core.hpp"

namespace vsmcx {

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
    // a select, not b.v[word]: a runtime index would send the block through
    // scratch memory
    return word == 0 ? b.v[0] : word == 1 ? b.v[1] : word == 2 ? b.v[2] : b.v[3];
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
VSMCX_HD double u01_oc_53(uint32_t u) { return u01_open_closed_32_53(u); }

// u01_open_open_32_24 (include/vsmc/rng/u01.h:116-118): (0.5 + (u >> 9)) 2^-23,
// the midpoint of one of 2^23 equal cells, so its mean is that of the 32-bit
// uniform.  The open-closed 24-bit form (u01.h:112-113) takes the top of each
// 2^-24 cell instead, which lowers E[-2 log u] by about 1e-6: every FP32
// normal from it would be slightly short, a bias shared by all particles.
// The kernels take the same value from its bits (u01_oo_24_bits: two
// instructions instead of a shift, a conversion, an add and a multiply).
VSMCX_HD float u01_oo_24(uint32_t u) { return u01_oo_24_bits(u); }

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
