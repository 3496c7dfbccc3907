// device_util.hpp -- wave64 / workgroup primitives for gfx950.
//
// Cross-lane moves are overloads per type (field by field for structs), so
// every value stays in VGPRs: no struct is ever addressed through memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsmcx {

constexpr int kWave = 64;

// ---- cross-lane moves of scalars ------------------------------------------------
__device__ __forceinline__ double lane_up(double v, unsigned d) { return __shfl_up(v, d, kWave); }
__device__ __forceinline__ int lane_up(int v, unsigned d) { return __shfl_up(v, d, kWave); }
__device__ __forceinline__ uint32_t lane_up(uint32_t v, unsigned d)
{
    return (uint32_t)__shfl_up((int)v, d, kWave);
}
__device__ __forceinline__ uint64_t lane_up(uint64_t v, unsigned d)
{
    return (uint64_t)__shfl_up((unsigned long long)v, d, kWave);
}
__device__ __forceinline__ double lane_get(double v, int l) { return __shfl(v, l, kWave); }
__device__ __forceinline__ int lane_get(int v, int l) { return __shfl(v, l, kWave); }
__device__ __forceinline__ uint32_t lane_get(uint32_t v, int l) { return (uint32_t)__shfl((int)v, l, kWave); }
__device__ __forceinline__ uint64_t lane_get(uint64_t v, int l)
{
    return (uint64_t)__shfl((unsigned long long)v, l, kWave);
}
__device__ __forceinline__ double lane_xor(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ int lane_xor(int v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ uint64_t lane_xor(uint64_t v, int m)
{
    return (uint64_t)__shfl_xor((unsigned long long)v, m, kWave);
}

// Inclusive scan across the 64 lanes of a wave (Hillis-Steele; op(earlier, later)).
template <typename T, typename Op>
__device__ __forceinline__ T wave_incl_scan(T v, Op op)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T o = lane_up(v, d);
        if (lane >= d) v = op(o, v);
    }
    return v;
}

// Reduction across the wave; every lane gets the result.
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op)
{
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) v = op(v, lane_xor(v, m));
    return v;
}

// Inclusive scan over a workgroup of BLOCK threads.  `lds` must hold BLOCK/64
// elements and keeps the inclusive per-wave totals afterwards.  `total`
// (optional) receives the workgroup aggregate.
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
            T o = lane_up(w, d);
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

// Exclusive counterpart: returns the combination of all earlier threads
// (identity for thread 0).
template <int BLOCK, typename T, typename Op>
__device__ __forceinline__ T block_excl_scan(T v, Op op, T ident, T *lds, T *total)
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    T incl = block_incl_scan<BLOCK>(v, op, lds, total);
    T up = lane_up(incl, 1);
    if (lane == 0) up = wid > 0 ? lds[wid - 1] : ident;
    __syncthreads();
    return up;
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
struct MinOp {
    template <typename T>
    __device__ __forceinline__ T operator()(const T &a, const T &b) const { return b < a ? b : a; }
};

// Look-back status words.  Polls are relaxed agent-scope loads (coherent at
// the device level, but without the L1 invalidation an acquire load carries);
// a reader issues one acquire fence after it has seen every flag it needs and
// before it reads the payloads those flags publish.
__device__ __forceinline__ uint64_t atomic_load_acq(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_load_rlx(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acq_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void atomic_store_rel(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace vsmcx
