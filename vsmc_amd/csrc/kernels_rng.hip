// kernels_rng.hip -- bulk counter-based RNG blocks.
//
// Block k = engine block at counter ctr0 + 1 + k (the counter is incremented
// before each block, include/vsmc/rng/philox.hpp:386-395).  One thread per
// block, each writing 4 words; stores are 16 B (Philox/Threefry4x32) or 32 B
// (Threefry4x64) per lane so a wave writes one contiguous 1-2 KiB span.
#include "internal.hpp"
#include "launch.hpp"
#include "rng_core.hpp"

namespace vsmcx {

__global__ __launch_bounds__(256) void k_philox4x32_fill(uint32_t k0, uint32_t k1, u32x4 c0,
                                                         uint64_t nblocks, uint4 *out)
{
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nblocks;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const u32x4 b = philox4x32_10(ctr_add(c0, k + 1), k0, k1);
        out[k] = make_uint4(b.v[0], b.v[1], b.v[2], b.v[3]);
    }
}

struct Key64 { uint64_t k[4]; };
struct Key32 { uint32_t k[4]; };

__global__ __launch_bounds__(256) void k_threefry4x64_fill(Key64 key, u64x4 c0, uint64_t nblocks,
                                                           ulonglong2 *out)
{
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nblocks;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const u64x4 b = threefry4x64_20(ctr_add(c0, k + 1), key.k);
        out[2 * k] = make_ulonglong2(b.v[0], b.v[1]);
        out[2 * k + 1] = make_ulonglong2(b.v[2], b.v[3]);
    }
}

__global__ __launch_bounds__(256) void k_threefry4x32_fill(Key32 key, u32x4 c0, uint64_t nblocks,
                                                           uint4 *out)
{
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nblocks;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const u32x4 b = threefry4x32_20(ctr_add(c0, k + 1), key.k);
        out[k] = make_uint4(b.v[0], b.v[1], b.v[2], b.v[3]);
    }
}

static unsigned grid_for(uint64_t n)
{
    const uint64_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

void launch_philox4x32_fill(uint32_t k0, uint32_t k1, const uint32_t ctr0[4], uint64_t nblocks,
                            uint32_t *out, hipStream_t st)
{
    u32x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr0[j];
    hipLaunchKernelGGL(k_philox4x32_fill, dim3(grid_for(nblocks)), dim3(256), 0, st, k0, k1, c,
                       nblocks, reinterpret_cast<uint4 *>(out));
}

void launch_threefry4x64_fill(const uint64_t key[4], const uint64_t ctr0[4], uint64_t nblocks,
                              uint64_t *out, hipStream_t st)
{
    u64x4 c;
    Key64 k;
    for (int j = 0; j < 4; ++j) {
        c.v[j] = ctr0[j];
        k.k[j] = key[j];
    }
    hipLaunchKernelGGL(k_threefry4x64_fill, dim3(grid_for(nblocks)), dim3(256), 0, st, k, c,
                       nblocks, reinterpret_cast<ulonglong2 *>(out));
}

void launch_threefry4x32_fill(const uint32_t key[4], const uint32_t ctr0[4], uint64_t nblocks,
                              uint32_t *out, hipStream_t st)
{
    u32x4 c;
    Key32 k;
    for (int j = 0; j < 4; ++j) {
        c.v[j] = ctr0[j];
        k.k[j] = key[j];
    }
    hipLaunchKernelGGL(k_threefry4x32_fill, dim3(grid_for(nblocks)), dim3(256), 0, st, k, c,
                       nblocks, reinterpret_cast<uint4 *>(out));
}

}  // namespace vsmcx
