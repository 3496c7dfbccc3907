// vsmcx::PfSampler run as the pf example runs vSMC's Sampler
// (example/pf/include/pf_mpi_do.hpp:36-64): Sampler(N, scheme, 0.5),
// initialize(data file), iterate(T - 1), then `out << sampler << std::endl`.
// The table goes to stdout.
//
// Usage: pf_sampler data_file N scheme fp64 seed resample_seed threshold
//
// Test infrastructure: built by tests/cpp/Makefile against include/ and
// vsmc_amd/libvsmcx.so, run on a GPU by tests/test_gpu_sampler.py.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "vsmcx.hpp"

int main(int argc, char **argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: pf_sampler data_file N scheme fp64 seed resample_seed threshold\n");
        return 2;
    }
    const std::size_t N = std::strtoull(argv[2], nullptr, 10);
    vsmcx::FilterConfig cfg;
    cfg.precision = std::atoi(argv[4]) ? VSMCX_FP64_FUSED : VSMCX_MIXED;
    cfg.seed = std::strtoull(argv[5], nullptr, 10);
    cfg.resample_seed = std::strtoull(argv[6], nullptr, 10);
    std::size_t T = 0;
    {
        std::ifstream in(argv[1]);
        double x, y;
        while (in >> x >> y) ++T;
    }
    try {
        vsmcx::PfSampler sampler(vsmcx::default_context(), N,
                                 (vsmcx::ResampleScheme)std::atoi(argv[3]), std::atof(argv[7]),
                                 cfg);
        sampler.initialize(argv[1]);
        sampler.iterate(T - 1);
        std::cout << sampler << std::endl;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
