// state.hpp -- the device-resident particle values behind vsmcx_state_*
// (state.cpp), shared with the device weight set (weights.cpp), which gathers
// them after a resample.
#pragma once
#include <hip/hip_runtime.h>

#include "host.hpp"

// Rows are n x dim values, FP32 (mixed) or FP64, kept in two device buffers:
// a copy gathers into the other one and swaps.
struct vsmcx_state {
    vsmcx_ctx *ctx = nullptr;
    uint64_t n = 0;
    uint32_t dim = 0;
    bool fp64 = false;
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;
    uint64_t *cf = nullptr;  // copy_from on the device
    double *inc = nullptr;   // increments of the last move
    int *err = nullptr;

    size_t rb() const { return fp64 ? 8 : 4; }
    size_t bytes() const { return rb() * n * dim; }

    void release_all()
    {
        void *ps[] = {buf[0], buf[1], cf, inc, err};
        for (void *p : ps) ctx->release(p);
    }

    void check()
    {
        int e = 0;
        VSMCX_HIP(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (e) throw ArgError("copy_from holds an index outside the particle range");
    }
};

