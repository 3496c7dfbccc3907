// lg_model.hpp -- the linear-Gaussian model of the serving workload:
// x_t = A x_{t-1} + q z, y_t = x_t + r v, x_0 ~ N(0, s0^2 I), dimension <= 10.
// The normals of particle i at step t come from Philox4x32-10 blocks
// {t+1, id_lo, id_hi, j}, j = 0, 1, 2 (four normals each, Box-Muller).
#pragma once
#include <math.h>

#include "internal.hpp"
#include "launch.hpp"
#include "rng_core.hpp"

namespace vsmcx {

// four N(0,1) from Philox block {t+1, id_lo, id_hi, j} (Box-Muller, as cv_normals)
__device__ __forceinline__ void lg_normals(float z[4], uint32_t k0, uint32_t k1, uint64_t id,
                                           uint64_t t, uint32_t j)
{
    u32x4 c;
    c.v[0] = (uint32_t)(t + 1);
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = j;
    const u32x4 w = philox4x32_10(c, k0, k1);
    box_muller_f32(w.v[0], w.v[1], z[0], z[1]);
    box_muller_f32(w.v[2], w.v[3], z[2], z[3]);
}

__device__ __forceinline__ void lg_normals(double z[4], uint32_t k0, uint32_t k1, uint64_t id,
                                           uint64_t t, uint32_t j)
{
    u32x4 c;
    c.v[0] = (uint32_t)(t + 1);
    c.v[1] = (uint32_t)id;
    c.v[2] = (uint32_t)(id >> 32);
    c.v[3] = j;
    const u32x4 w = philox4x32_10(c, k0, k1);
    const double two_pi = 6.2831853071795865;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const double u1 = u01_oc_53(w.v[2 * h]);
        const double u2 = u01_oc_53(w.v[2 * h + 1]);
        const double r = sqrt(-2 * log(u1));
        z[2 * h] = r * cos(two_pi * u2);
        z[2 * h + 1] = r * sin(two_pi * u2);
    }
}

template <typename Real>
__device__ __forceinline__ Real lg_a(const LgParams &lp, int k);
template <>
__device__ __forceinline__ float lg_a<float>(const LgParams &lp, int k) { return lp.a[k]; }
template <>
__device__ __forceinline__ double lg_a<double>(const LgParams &lp, int k) { return lp.ad[k]; }

}  // namespace vsmcx
