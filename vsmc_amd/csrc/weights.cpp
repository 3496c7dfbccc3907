// weights.cpp -- the device weight set (vsmcx_weights_*): vSMC's WeightSet
// with its log-weights in HBM, fed by device arrays of increments (a user
// kernel's move), and Particle::resample on the device.
//
// Host code only.  A set or add is one tile pass (k_ws_add) and the filter's
// one-wave reduction (ESS, log-evidence, the resample decision); a resample
// is the filter's exact-fold pipeline on these log-weights, then the state's
// gather in place.  Nothing waits for the device until the history is read.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "state.hpp"

using namespace vsmcx;
using namespace vsmcx_host;

struct vsmcx_weights {
    vsmcx_ctx *ctx = nullptr;
    uint64_t n = 0, tiles = 0, max_steps = 0;
    int scheme = 0;
    uint64_t t = 0;        // history row of the last set / add
    bool started = false;  // a set_log has been made
    double *lw = nullptr, *inc = nullptr, *part = nullptr;
    FilterCtl *ctl = nullptr;
    StepRow *hist = nullptr;
    int *err = nullptr;
    ResampleBuffers rb;

    void release_all()
    {
        void *ps[] = {lw, inc, part, ctl, hist, err};
        for (void *p : ps) ctx->release(p);
        free_resample(ctx, rb);
    }
};

int vsmcx_ctx_stream(vsmcx_ctx *ctx, void **hip_stream)
{
    API_BEGIN
    REQUIRE(ctx && hip_stream, "NULL argument");
    *hip_stream = (void *)ctx->stream;
    API_END
}

int vsmcx_weights_create(vsmcx_ctx *ctx, uint64_t n, int scheme, double threshold,
                         uint64_t resample_seed, uint64_t max_steps, vsmcx_weights **out)
{
    API_BEGIN
    REQUIRE(ctx && out, "NULL argument");
    REQUIRE(n >= 1 && n < (1ull << 31), "n must be in [1, 2^31)");
    REQUIRE(scheme_ok(scheme), "unknown resampling scheme");
    REQUIRE(max_steps >= 1, "max_steps must be >= 1");
    std::unique_ptr<vsmcx_weights> W(new vsmcx_weights);
    W->ctx = ctx;
    W->n = n;
    W->tiles = (n + kMoveTile - 1) / kMoveTile;
    W->max_steps = max_steps;
    W->scheme = scheme;
    try {
        W->lw = ctx->alloc_n<double>(n);
        W->inc = ctx->alloc_n<double>(n);
        W->part = ctx->alloc_n<double>(5 * W->tiles);
        W->ctl = ctx->alloc_n<FilterCtl>(1);
        W->hist = ctx->alloc_zero<StepRow>(max_steps);
        W->err = ctx->alloc_zero<int>(1);
        alloc_resample(ctx, W->rb, n, n, true);
    } catch (...) {
        W->release_all();
        throw;
    }
    FilterCtl h{};
    h.n = n;
    h.threshold = threshold;
    h.scheme = scheme;
    h.eng.key[0] = resample_seed;  // Threefry4x64 seed(s): key {s, 0, 0, 0}, buffer empty
    h.eng.index = 4;
    VSMCX_HIP(hipMemcpyAsync(W->ctl, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    *out = W.release();
    API_END
}

int vsmcx_weights_destroy(vsmcx_weights *w)
{
    API_BEGIN
    if (!w) return VSMCX_OK;
    (void)hipStreamSynchronize(w->ctx->stream);
    w->release_all();
    delete w;
    API_END
}

int vsmcx_weights_buffer(vsmcx_weights *w, double **inc_dev)
{
    API_BEGIN
    REQUIRE(w && inc_dev, "NULL argument");
    *inc_dev = w->inc;
    API_END
}

static void ws_update(vsmcx_weights *w, const double *inc_dev, bool set)
{
    REQUIRE(inc_dev, "NULL increments");
    if (set) {
        w->t = 0;
        w->started = true;
    } else {
        if (!w->started) throw DevError(VSMCX_ERR_STATE, "add_log before set_log");
        if (w->t + 1 >= w->max_steps) throw DevError(VSMCX_ERR_STATE, "history full (max_steps)");
        ++w->t;
    }
    hipStream_t s = w->ctx->stream;
    launch_ws_add(w->ctl, w->lw, inc_dev, w->n, set ? 1 : 0, w->part, s);
    launch_cv_reduce(w->ctl, w->part, w->tiles, w->hist, w->t, s);
    VSMCX_HIP(hipGetLastError());
}

int vsmcx_weights_set_log(vsmcx_weights *w, const double *inc_dev)
{
    API_BEGIN
    REQUIRE(w, "NULL weight set");
    ws_update(w, inc_dev, true);
    API_END
}

int vsmcx_weights_add_log(vsmcx_weights *w, const double *inc_dev)
{
    API_BEGIN
    REQUIRE(w, "NULL weight set");
    ws_update(w, inc_dev, false);
    API_END
}

int vsmcx_weights_resample(vsmcx_weights *w, vsmcx_state *state)
{
    API_BEGIN
    REQUIRE(w, "NULL weight set");
    REQUIRE(w->started, "resample before set_log");
    REQUIRE(!state || state->n == w->n, "the state must hold one row per particle");
    hipStream_t s = w->ctx->stream;
    launch_resample_from_log_weights(w->ctl, w->lw, w->n, w->scheme, w->rb, w->hist, s, w->part);
    if (state)
        launch_ws_gather(state->fp64, w->ctl, state->buf[state->cur], w->rb.anc, state->dim, w->n,
                         w->err, s);
    launch_ws_done(w->ctl, s);
    VSMCX_HIP(hipGetLastError());
    API_END
}

int vsmcx_weights_ancestors(vsmcx_weights *w, const uint32_t **anc_dev)
{
    API_BEGIN
    REQUIRE(w && anc_dev, "NULL argument");
    *anc_dev = w->rb.anc;
    API_END
}

int vsmcx_weights_steps(vsmcx_weights *w, uint64_t *steps)
{
    API_BEGIN
    REQUIRE(w && steps, "NULL argument");
    *steps = w->started ? w->t + 1 : 0;
    API_END
}

int vsmcx_weights_history(vsmcx_weights *w, uint64_t count, vsmcx_step_stats *out)
{
    API_BEGIN
    REQUIRE(w && (out || count == 0), "NULL argument");
    REQUIRE(count <= (w->started ? w->t + 1 : 0), "count exceeds the steps taken");
    hipStream_t s = w->ctx->stream;
    FilterCtl h;
    int e = 0;
    VSMCX_HIP(hipMemcpyAsync(&h, w->ctl, sizeof(h), hipMemcpyDeviceToHost, s));
    VSMCX_HIP(hipMemcpyAsync(&e, w->err, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<StepRow> rows(count);
    if (count)
        VSMCX_HIP(hipMemcpyAsync(rows.data(), w->hist, count * sizeof(StepRow),
                                 hipMemcpyDeviceToHost, s));
    w->ctx->sync();
    check_dev_err(h);
    if (e) throw DevError(VSMCX_ERR_INTERNAL, "ancestor outside the particle range");
    for (uint64_t i = 0; i < count; ++i) {
        out[i].ess = rows[i].ess;
        out[i].log_z = rows[i].logz;
        out[i].mean_x = rows[i].mean_x;
        out[i].mean_y = rows[i].mean_y;
        out[i].resampled = rows[i].resampled;
        out[i].pad = 0;
        out[i].draws = rows[i].draws;
    }
    API_END
}

int vsmcx_weights_read(vsmcx_weights *w, double *weight)
{
    API_BEGIN
    REQUIRE(w && weight, "NULL argument");
    REQUIRE(w->started, "read before set_log");
    hipStream_t s = w->ctx->stream;
    FilterCtl h;
    VSMCX_HIP(hipMemcpyAsync(&h, w->ctl, sizeof(h), hipMemcpyDeviceToHost, s));
    w->ctx->sync();
    if (h.prev_resampled) {  // set_equal_weight after the resample
        for (uint64_t i = 0; i < w->n; ++i) weight[i] = 1.0 / (double)w->n;
        return VSMCX_OK;
    }
    // the fold's x scratch (n doubles) holds them on the way
    launch_cv_weights(w->ctl, w->lw, w->n, w->rb.fs.wx, s);
    VSMCX_HIP(hipMemcpyAsync(weight, w->rb.fs.wx, 8 * w->n, hipMemcpyDeviceToHost, s));
    w->ctx->sync();
    API_END
}
