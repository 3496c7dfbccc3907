// device_util.hpp -- wave64 / workgroup primitives for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsmcx {

constexpr int kWave = 64;

// Cross-lane moves of any 4-byte-multiple POD, word by word (ds_bpermute).
template <typename T>
__device__ __forceinline__ T shfl_up_t(const T &v, unsigned d)
{
    static_assert(sizeof(T) % 4 == 0, "shfl of non-word type");
    T r;
    const int *s = reinterpret_cast<const int *>(&v);
    int *o = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __shfl_up(s[i], d, kWave);
    return r;
}

template <typename T>
__device__ __forceinline__ T shfl_t(const T &v, int lane)
{
    static_assert(sizeof(T) % 4 == 0, "shfl of non-word type");
    T r;
    const int *s = reinterpret_cast<const int *>(&v);
    int *o = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __shfl(s[i], lane, kWave);
    return r;
}

template <typename T>
__device__ __forceinline__ T shfl_xor_t(const T &v, int m)
{
    static_assert(sizeof(T) % 4 == 0, "shfl of non-word type");
    T r;
    const int *s = reinterpret_cast<const int *>(&v);
    int *o = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __shfl_xor(s[i], m, kWave);
    return r;
}

// Inclusive scan across the 64 lanes of a wave (Hillis-Steele; op(earlier, later)).
template <typename T, typename Op>
__device__ __forceinline__ T wave_incl_scan(T v, Op op)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T o = shfl_up_t(v, d);
        if (lane >= d) v = op(o, v);
    }
    return v;
}

// Reduction across the wave; every lane gets the result.
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op)
{
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) v = op(v, shfl_xor_t(v, m));
    return v;
}

// Inclusive scan over a workgroup of BLOCK threads.  `lds` must hold BLOCK/64
// elements.  `total` (optional) receives the workgroup aggregate.
template <int BLOCK, typename T, typename Op>
__device__ __forceinline__ T block_incl_scan(T v, Op op, T *lds, T *total)
{
    constexpr int NW = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_incl_scan(v, op);
    if (lane == kWave - 1) lds[wid] = v;
    __syncthreads();
    if (wid == 0) {
        T w = lds[lane < NW ? lane : NW - 1];
#pragma unroll
        for (int d = 1; d < NW; d <<= 1) {
            T o = shfl_up_t(w, d);
            if (lane >= d) w = op(o, w);
        }
        if (lane < NW) lds[lane] = w;
    }
    __syncthreads();
    if (wid > 0) v = op(lds[wid - 1], v);
    if (total) *total = lds[NW - 1];
    __syncthreads();
    return v;
}

template <int BLOCK, typename T, typename Op>
__device__ __forceinline__ T block_incl_scan(T v, Op op, T *lds)
{
    return block_incl_scan<BLOCK>(v, op, lds, (T *)nullptr);
}

template <int BLOCK, typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T *lds)
{
    constexpr int NW = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_reduce(v, op);
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    T r = lds[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = op(r, lds[i]);
    __syncthreads();
    return r;
}

struct AddOp {
    template <typename T>
    __device__ __forceinline__ T operator()(const T &a, const T &b) const { return a + b; }
};
struct MaxOp {
    template <typename T>
    __device__ __forceinline__ T operator()(const T &a, const T &b) const { return a < b ? b : a; }
};

// Relaxed/acquire-release helpers for decoupled look-back status words.
__device__ __forceinline__ uint64_t atomic_load_acq(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_rel(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace vsmcx
