// Drop-in check of include/vsmcx.hpp against vSMC itself: two
// vsmc::Particle objects with the same log-weights and the same resample
// engine state are resampled, one with vSMC's own Resample<Scheme> operator
// (include/vsmc/resample/*.hpp:53-77) and one with vsmcx::HipResample, through
// the unmodified Particle::resample (include/vsmc/core/particle.hpp:230-250).
// The particles kept (state = original index, so state after the copy = the
// ancestor) and the resample engines afterwards must be identical.
//
// Test infrastructure: compiled by oracle/Makefile (target dropin) against the
// reference headers into oracle/_ref/, linked with vsmc_amd/libvsmcx.so, and
// run on a GPU by tests/test_gpu_dropin.py.
#include <cmath>
#include <cstdio>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <vsmc/core/particle.hpp>
#include <vsmc/core/state_matrix.hpp>
#include <vsmc/resample/multinomial.hpp>
#include <vsmc/resample/residual.hpp>
#include <vsmc/resample/residual_stratified.hpp>
#include <vsmc/resample/residual_systematic.hpp>
#include <vsmc/resample/stratified.hpp>
#include <vsmc/resample/systematic.hpp>

#include "vsmcx.hpp"

typedef vsmc::StateMatrix<vsmc::RowMajor, 1, double> Value;
typedef vsmc::Particle<Value> Part;

static Part::resample_type vsmc_op(int scheme)
{
    switch (scheme) {
    case 0: return vsmc::ResampleType<vsmc::Multinomial>::type();
    case 1: return vsmc::ResampleType<vsmc::Residual>::type();
    case 2: return vsmc::ResampleType<vsmc::Stratified>::type();
    case 3: return vsmc::ResampleType<vsmc::Systematic>::type();
    case 4: return vsmc::ResampleType<vsmc::ResidualStratified>::type();
    default: return vsmc::ResampleType<vsmc::ResidualSystematic>::type();
    }
}

static std::string engine_text(Part &p)
{
    std::stringstream ss;
    ss << p.resample_rng();
    return ss.str();
}

int main()
{
    vsmcx::Context ctx(0);
    std::mt19937_64 g(2024);
    std::normal_distribution<double> Z(0, 1);
    int failures = 0, cases = 0;
    const std::size_t sizes[] = {1, 7, 1000, 65537, 1u << 20};
    for (int scheme = 0; scheme < 6; ++scheme) {
        vsmcx::HipResample hip(ctx, (vsmcx::ResampleScheme)scheme);
        for (std::size_t N : sizes) {
            Part a(N), b(N);
            b.resample_rng() = a.resample_rng();
            std::vector<double> lw(N);
            for (int rep = 0; rep < 3; ++rep) {
                const double spread = rep == 0 ? 1.0 : (rep == 1 ? 5.0 : 30.0);
                for (std::size_t i = 0; i < N; ++i) {
                    lw[i] = -0.5 * std::pow(spread * Z(g), 2);
                    a.value().state(i, 0) = (double)i;
                    b.value().state(i, 0) = (double)i;
                }
                a.weight_set().set_log_weight(lw.begin());
                b.weight_set().set_log_weight(lw.begin());
                const bool ra = a.resample(vsmc_op(scheme), 2.0);  // always resample
                const bool rb = b.resample(hip, 2.0);
                std::size_t bad = 0;
                for (std::size_t i = 0; i < N; ++i)
                    bad += a.value().state(i, 0) != b.value().state(i, 0);
                const bool same_eng = engine_text(a) == engine_text(b);
                ++cases;
                if (!ra || !rb || bad || !same_eng) {
                    ++failures;
                    std::printf("MISMATCH scheme %d N %zu rep %d: %zu ancestors differ, engine %s\n",
                                scheme, N, rep, bad, same_eng ? "same" : "differs");
                }
            }
        }
        std::printf("scheme %d done\n", scheme);
        std::fflush(stdout);
    }
    std::printf("dropin cases %d failures %d\n", cases, failures);
    return failures ? 1 : 0;
}
