// The public device header (include/vsmcx_device.hpp) compiled for the host,
// with C entry points for the CPU tests (tests/test_device_header.py): its
// block functions, urng streams, normal01 streams and u01 conversions.
//
// Test infrastructure: built by tests/cpp/Makefile into
// tests/cpp/_bin/libvsmcx_device_host.so.
#include <cstddef>
#include <cstdint>

#include "vsmcx_device.hpp"

using namespace vsmcx;

extern "C" {

void dh_philox4x32_block(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    u32x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr[j];
    const u32x4 r = philox4x32_10(c, key[0], key[1]);
    for (int j = 0; j < 4; ++j) out[j] = r.v[j];
}

void dh_threefry4x32_block(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    u32x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr[j];
    const u32x4 r = threefry4x32_20(c, key);
    for (int j = 0; j < 4; ++j) out[j] = r.v[j];
}

void dh_threefry4x64_block(const uint64_t *ctr, const uint64_t *key, uint64_t *out)
{
    u64x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr[j];
    const u64x4 r = threefry4x64_20(c, key);
    for (int j = 0; j < 4; ++j) out[j] = r.v[j];
}

// kind 0 philox4x32, 1 threefry4x32 (cburng4x32), 2 threefry4x64 (cburng4x64)
void dh_urng_stream(int kind, uint64_t seed, size_t n, uint64_t *out)
{
    philox4x32_rng_t p;
    threefry4x32_rng_t t32;
    threefry4x64_rng_t t64;
    philox4x32_init(&p, (uint32_t)seed);
    threefry4x32_init(&t32, (uint32_t)seed);
    threefry4x64_init(&t64, seed);
    for (size_t i = 0; i < n; ++i)
        out[i] = kind == 0 ? philox4x32_rand(&p)
                 : kind == 1 ? threefry4x32_rand(&t32)
                             : threefry4x64_rand(&t64);
}

// kind 0 normal01_4x32_24, 1 normal01_4x32_53, 2 normal01_4x64_53
void dh_normal01_stream(int kind, uint64_t seed, size_t n, double *out)
{
    cburng4x32_rng_t r32;
    cburng4x64_rng_t r64;
    cburng4x32_init(&r32, (uint32_t)seed);
    cburng4x64_init(&r64, seed);
    normal01_4x32_24 a{};
    normal01_4x32_53 b{};
    normal01_4x64_53 c{};
    if (kind == 0) normal01_4x32_24_init(&a, &r32);
    if (kind == 1) normal01_4x32_53_init(&b, &r32);
    if (kind == 2) normal01_4x64_53_init(&c, &r64);
    for (size_t i = 0; i < n; ++i)
        out[i] = kind == 0 ? (double)normal01_4x32_24_rand(&a, &r32)
                 : kind == 1 ? normal01_4x32_53_rand(&b, &r32)
                             : normal01_4x64_53_rand(&c, &r64);
}

// the layout of the reference harness's ref_u01_32 / ref_u01_64
void dh_u01_32(const uint32_t *u, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) {
        double *o = out + 8 * i;
        o[0] = u01_closed_closed_32_24(u[i]);
        o[1] = u01_closed_open_32_24(u[i]);
        o[2] = u01_open_closed_32_24(u[i]);
        o[3] = u01_open_open_32_24(u[i]);
        o[4] = u01_closed_closed_32_53(u[i]);
        o[5] = u01_closed_open_32_53(u[i]);
        o[6] = u01_open_closed_32_53(u[i]);
        o[7] = u01_open_open_32_53(u[i]);
    }
}

void dh_u01_64(const uint64_t *u, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) {
        double *o = out + 4 * i;
        o[0] = u01_closed_closed_64_53(u[i]);
        o[1] = u01_closed_open_64_53(u[i]);
        o[2] = u01_open_closed_64_53(u[i]);
        o[3] = u01_open_open_64_53(u[i]);
    }
}

}  // extern "C"
