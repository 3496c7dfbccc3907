// kernels_probe.hip -- streaming rates of the GPU: a plain copy, a read and a
// write of one large buffer, 16 B per lane per access, grid-stride.  They give
// the HBM rate that simple streaming reaches on this part, the yardstick for
// the particle kernels' achieved rates next to the published peak.
#include "launch.hpp"

namespace vsmcx {

__global__ __launch_bounds__(256) void k_probe_copy(const float4 *__restrict__ a,
                                                    float4 *__restrict__ b, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// The sum keeps the loads live; it is stored only when it equals a value the
// probe's zero-filled input cannot produce.
__global__ __launch_bounds__(256) void k_probe_read(const float4 *__restrict__ a,
                                                    float4 *__restrict__ out, uint64_t n)
{
    float acc = 0.f;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    if (acc == -1.f) out[0] = make_float4(acc, acc, acc, acc);
}

__global__ __launch_bounds__(256) void k_probe_write(float4 *__restrict__ b, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i] = make_float4(0.f, 1.f, 2.f, (float)(uint32_t)i);
}

// Best time in ms of `reps` launches of probe `kind` (0 copy a -> b, 1 read a,
// 2 write b) over n float4 elements, over workgroup counts 1024 and 4096,
// timed by HIP events on st.  Returns a negative value on a HIP error.
float launch_stream_probe(int kind, const float4 *a, float4 *b, uint64_t n, int reps,
                          hipStream_t st)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1.f;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -1.f;
    }
    float best = -1.f;
    bool ok = true;
    for (unsigned grid : {1024u, 4096u}) {
        for (int r = 0; ok && r <= reps; ++r) {  // the first launch of a shape is not timed
            ok = hipEventRecord(e0, st) == hipSuccess;
            if (!ok) break;
            if (kind == 0)
                hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, st, a, b, n);
            else if (kind == 1)
                hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, st, a, b, n);
            else
                hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(256), 0, st, b, n);
            float ms = 0.f;
            ok = hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok && r > 0 && (best < 0.f || ms < best)) best = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ok ? best : -1.f;
}

}  // namespace vsmcx
