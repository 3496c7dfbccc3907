// Drop-in check of vsmcx::StateHIP and vsmcx::CvMoveHIP against vSMC itself:
// an unmodified vsmc::Particle<vsmcx::StateHIP<4, FP>> runs the pf model in
// vSMC's step order (core/sampler.hpp:773-834): init, then per iteration the
// move, Particle::resample (include/vsmc/core/particle.hpp:230-250) with
// vsmcx::HipResample, and the MonitorMCMC mean.  The particles live in HBM
// (StateHIP::copy gathers there); the weights live in vSMC's own host
// WeightSet.  One line per iteration goes to stdout:
//   t ess log_z mean_x mean_y resampled
// with log_z the running log sum_i W_i exp(inc_i) (example/include/
// move_smc.hpp:64-71, max-shifted).
//
// Usage: dropin_state N T scheme fp64 seed resample_seed threshold < obs
// (T lines "x y").
//
// Test infrastructure: compiled by oracle/Makefile (target dropin) against the
// reference headers into oracle/_ref/, linked with vsmc_amd/libvsmcx.so, run
// on a GPU by tests/test_gpu_dropin.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <vsmc/core/particle.hpp>

#include "vsmcx.hpp"

template <typename FP>
static int run(std::size_t N, std::size_t T, int scheme, std::uint64_t seed, std::uint64_t rseed,
               double threshold, const std::vector<double> &obs)
{
    typedef vsmcx::StateHIP<4, FP> Value;
    vsmc::Particle<Value> p(N);
    p.resample_rng().seed(rseed);
    p.weight_set().set_equal_weight();  // Sampler::do_reset (core/sampler.hpp:758-771)
    vsmcx::FilterConfig cfg;
    cfg.n = N;
    cfg.horizon = T;
    cfg.precision = sizeof(FP) == 8 ? VSMCX_FP64_FUSED : VSMCX_MIXED;
    cfg.seed = seed;
    vsmcx::CvMoveHIP move(cfg, obs);
    vsmcx::HipResample op(vsmcx::default_context(), (vsmcx::ResampleScheme)scheme);
    std::vector<double> w(N);
    double log_z = 0;
    for (std::size_t t = 0; t < T; ++t) {
        p.weight_set().read_weight(w.begin());
        if (t == 0)
            move(p, nullptr);
        else
            move(t, p);
        const std::vector<double> &inc = move.increments();
        double imax = inc[0];
        for (std::size_t i = 0; i < N; ++i) imax = inc[i] > imax ? inc[i] : imax;
        double sum = 0;
        for (std::size_t i = 0; i < N; ++i) sum += w[i] * std::exp(inc[i] - imax);
        log_z += std::log(sum) + imax;
        const double ess = p.weight_set().ess();
        const bool resampled = p.resample(op, threshold);
        p.weight_set().read_weight(w.begin());
        const std::vector<double> rows = p.value().rows();
        double mx = 0, my = 0;
        for (std::size_t i = 0; i < N; ++i) {
            mx += w[i] * rows[4 * i];
            my += w[i] * rows[4 * i + 1];
        }
        std::printf("%zu %.17g %.17g %.17g %.17g %d\n", t, ess, log_z, mx, my, resampled ? 1 : 0);
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: dropin_state N T scheme fp64 seed resample_seed threshold < obs\n");
        return 2;
    }
    const std::size_t N = std::strtoull(argv[1], nullptr, 10), T = std::strtoull(argv[2], nullptr, 10);
    const int scheme = std::atoi(argv[3]), fp64 = std::atoi(argv[4]);
    const std::uint64_t seed = std::strtoull(argv[5], nullptr, 10);
    const std::uint64_t rseed = std::strtoull(argv[6], nullptr, 10);
    const double threshold = std::atof(argv[7]);
    std::vector<double> obs(2 * T);
    for (std::size_t k = 0; k < 2 * T; ++k)
        if (std::scanf("%lf", &obs[k]) != 1) {
            std::fprintf(stderr, "short observation input\n");
            return 2;
        }
    try {
        return fp64 ? run<double>(N, T, scheme, seed, rseed, threshold, obs)
                    : run<float>(N, T, scheme, seed, rseed, threshold, obs);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
