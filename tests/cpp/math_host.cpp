// vsmc_amd/csrc/math_core.hpp compiled for the host, with C entry points for
// the CPU tests (tests/test_math_core.py): the FP64 exp of the weights and
// log-sum-exp partials, and the FP32 open-open uniform of the normals.  The
// host build runs the same IEEE operations (fma, rint, ldexp, a bit
// construction) as the device build.
//
// Test infrastructure: built by tests/cpp/Makefile into
// tests/cpp/_bin/libvsmcx_math_host.so.
#include <cstddef>
#include <cstdint>

#include "../../vsmc_amd/csrc/math_core.hpp"

extern "C" {

void mh_exp_lse(const double *d, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = vsmcx::exp_lse(d[i]);
}

void mh_u01_oo_24(const uint32_t *u, size_t n, float *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = vsmcx::u01_oo_24_bits(u[i]);
}

}  // extern "C"
