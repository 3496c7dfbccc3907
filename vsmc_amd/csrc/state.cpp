// state.cpp -- device-resident particle values for vSMC's Particle<T>
// (vsmcx_state_*): the value type's storage, its copy (StateMatrix::copy,
// include/vsmc/core/state_matrix.hpp:285-300) as a device gather, host
// access, and the pf model's move over such a state.
//
// Host code only.  Rows are n x dim values, FP32 (mixed) or FP64, kept in two
// device buffers: a copy gathers into the other one and swaps.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "state.hpp"

using namespace vsmcx;
using namespace vsmcx_host;

static vsmcx_state *state_new(vsmcx_ctx *ctx, uint64_t n, uint32_t dim, bool fp64)
{
    std::unique_ptr<vsmcx_state> S(new vsmcx_state);
    S->ctx = ctx;
    S->n = n;
    S->dim = dim;
    S->fp64 = fp64;
    try {
        S->buf[0] = ctx->alloc(S->bytes());
        S->buf[1] = ctx->alloc(S->bytes());
        S->cf = ctx->alloc_n<uint64_t>(n);
        S->inc = ctx->alloc_n<double>(n);
        S->err = ctx->alloc_zero<int>(1);
        VSMCX_HIP(hipMemsetAsync(S->buf[0], 0, S->bytes(), ctx->stream));
        ctx->sync();
    } catch (...) {
        S->release_all();
        throw;
    }
    return S.release();
}

int vsmcx_state_create(vsmcx_ctx *ctx, uint64_t n, uint32_t dim, int precision,
                       vsmcx_state **out)
{
    API_BEGIN
    REQUIRE(ctx && out, "NULL argument");
    REQUIRE(n >= 1 && dim >= 1, "empty state");
    REQUIRE(precision == VSMCX_MIXED || precision == VSMCX_FP64_FUSED,
            "state rows are FP32 (VSMCX_MIXED) or FP64 (VSMCX_FP64_FUSED)");
    *out = state_new(ctx, n, dim, precision == VSMCX_FP64_FUSED);
    API_END
}

int vsmcx_state_destroy(vsmcx_state *s)
{
    API_BEGIN
    if (!s) return VSMCX_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    s->release_all();
    delete s;
    API_END
}

int vsmcx_state_clone(vsmcx_state *s, vsmcx_state **out)
{
    API_BEGIN
    REQUIRE(s && out, "NULL argument");
    vsmcx_state *c = state_new(s->ctx, s->n, s->dim, s->fp64);
    const hipError_t e =
        hipMemcpyAsync(c->buf[0], s->buf[s->cur], s->bytes(), hipMemcpyDeviceToDevice, s->ctx->stream);
    if (e == hipSuccess) s->ctx->sync();
    if (e != hipSuccess) {
        c->release_all();
        delete c;
        VSMCX_HIP(e);
    }
    *out = c;
    API_END
}

int vsmcx_state_size(vsmcx_state *s, uint64_t *n, uint32_t *dim)
{
    API_BEGIN
    REQUIRE(s, "NULL state");
    if (n) *n = s->n;
    if (dim) *dim = s->dim;
    API_END
}

int vsmcx_state_write(vsmcx_state *s, const double *rows)
{
    API_BEGIN
    REQUIRE(s && rows, "NULL argument");
    const uint64_t m = s->n * s->dim;
    if (s->fp64) {
        VSMCX_HIP(hipMemcpyAsync(s->buf[s->cur], rows, 8 * m, hipMemcpyHostToDevice, s->ctx->stream));
        s->ctx->sync();
    } else {
        std::vector<float> f(rows, rows + m);
        VSMCX_HIP(hipMemcpyAsync(s->buf[s->cur], f.data(), 4 * m, hipMemcpyHostToDevice, s->ctx->stream));
        s->ctx->sync();
    }
    API_END
}

int vsmcx_state_read(vsmcx_state *s, double *rows)
{
    API_BEGIN
    REQUIRE(s && rows, "NULL argument");
    const uint64_t m = s->n * s->dim;
    if (s->fp64) {
        VSMCX_HIP(hipMemcpyAsync(rows, s->buf[s->cur], 8 * m, hipMemcpyDeviceToHost, s->ctx->stream));
        s->ctx->sync();
    } else {
        std::vector<float> f(m);
        VSMCX_HIP(hipMemcpyAsync(f.data(), s->buf[s->cur], 4 * m, hipMemcpyDeviceToHost, s->ctx->stream));
        s->ctx->sync();
        for (uint64_t j = 0; j < m; ++j) rows[j] = f[j];
    }
    API_END
}

int vsmcx_state_data(vsmcx_state *s, void **device_rows)
{
    API_BEGIN
    REQUIRE(s && device_rows, "NULL argument");
    *device_rows = s->buf[s->cur];
    API_END
}

int vsmcx_state_copy(vsmcx_state *s, uint64_t n, const uint64_t *copy_from)
{
    API_BEGIN
    REQUIRE(s && copy_from, "NULL argument");
    REQUIRE(n == s->n, "copy_from must hold one index per particle");
    hipStream_t st = s->ctx->stream;
    VSMCX_HIP(hipMemcpyAsync(s->cf, copy_from, 8 * n, hipMemcpyHostToDevice, st));
    launch_state_gather(s->fp64, s->buf[s->cur], s->buf[1 - s->cur], s->cf, s->dim, n, s->err, st);
    VSMCX_HIP(hipGetLastError());
    s->check();
    s->cur = 1 - s->cur;
    API_END
}

int vsmcx_state_cv_move(vsmcx_state *s, const vsmcx_filter_desc *m, uint64_t iter,
                        const double *obs_xy, double *inc)
{
    API_BEGIN
    REQUIRE(s && m && obs_xy && inc, "NULL argument");
    REQUIRE(s->dim == 4, "the pf model's state has 4 coordinates (x, y, vx, vy)");
    const CvParams p{m->dt, m->sd_pos, m->sd_vel, m->prior_sd_pos, m->prior_sd_vel, m->obs_scale,
                     m->nu, m->llh_const};
    hipStream_t st = s->ctx->stream;
    launch_state_cv_move(s->fp64, s->buf[s->cur], s->inc, p, m->seed, iter, obs_xy[0], obs_xy[1],
                         s->n, st);
    VSMCX_HIP(hipGetLastError());
    VSMCX_HIP(hipMemcpyAsync(inc, s->inc, 8 * s->n, hipMemcpyDeviceToHost, st));
    s->ctx->sync();
    API_END
}
