// capi.cpp -- the C ABI (include/vsmcx.h) over the HIP kernels.
//
// Host code only: contexts with a device-memory budget, the device-resident
// filter driver (one stream, no host synchronisation inside a step), the
// synchronous compatibility entries, and error translation at the boundary.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "resample_layout.hpp"
#include "rng_core.hpp"


namespace vsmcx {

// the context's budgeted allocator in the shape layout_resample takes
struct CtxAlloc {
    vsmcx_ctx *c;
    template <class T>
    T *n(uint64_t count) { return c->alloc_n<T>(count); }
};

// Fold / resample scratch for up to M particles (and nuni uniforms); own_ufs:
// the uniforms' fold gets scratch of its own (it runs on a side stream).
void alloc_resample(vsmcx_ctx *c, ResampleBuffers &rb, uint64_t M, uint64_t nuni, bool lists,
                    bool own_ufs)
{
    CtxAlloc a{c};
    layout_resample(a, rb, M, nuni, lists, own_ufs);
}

void free_resample(vsmcx_ctx *c, ResampleBuffers &rb)
{
    void *ps[] = {rb.anc, rb.zl, rb.sl, rb.rep_out, rb.a_out, rb.ufs.tsum, rb.ufs.ts, rb.uscan,
                  rb.u_tmax, rb.u_tmaxex, rb.fs.tsum, rb.fs.ts, rb.fs.agg, rb.fs.pre, rb.fs.spec,
                  rb.fs.gval, rb.fs.rep32, rb.fs.rep8, rb.fs.zs, rb.fs.zsx, rb.fs.wx, rb.fs.thr,
                  rb.fs.thr_fb, rb.fs.cagg, rb.fs.cpre, rb.fs.tcnt};
    for (void *p : ps) c->release(p);
    if (rb.ufs.agg != rb.fs.agg) {  // the uniforms' own fold scratch
        void *us[] = {rb.ufs.agg, rb.ufs.pre, rb.ufs.cagg, rb.ufs.cpre, rb.ufs.tcnt,
                      rb.ufs.spec, rb.ufs.gval, rb.ufs.thr, rb.ufs.thr_fb, rb.uguess};
        for (void *p : us) c->release(p);
    }
    rb = ResampleBuffers();
}

// the side stream's objects (nothing is left running on it once the filter's
// stream is done: every resample waits for the side stream's join)
void free_side(vsmcx_ctx *c, SideFold &sd)
{
    if (sd.s) {
        (void)hipStreamSynchronize(sd.s);
        (void)hipStreamDestroy(sd.s);
    }
    if (sd.fork) (void)hipEventDestroy(sd.fork);
    if (sd.join) (void)hipEventDestroy(sd.join);
    if (sd.starts) (void)hipEventDestroy(sd.starts);
    c->release(sd.uctl);
    sd = SideFold();
}

}  // namespace vsmcx

namespace {

Tf64State engine_in(const vsmcx_threefry4x64_state &e)
{
    Tf64State s{};
    for (int j = 0; j < 4; ++j) {
        s.ctr[j] = e.ctr[j];
        s.key[j] = e.key[j];
    }
    s.index = (uint32_t)(e.index > 4 ? 4 : e.index);
    return s;
}

// Engine state back in vSMC's layout.  vSMC's operator() leaves the block of
// the current counter in the buffer after every draw, also when the index has
// reached 4 (include/vsmc/rng/threefry.hpp:355-364); with no draw at all the
// buffer is whatever the caller's engine held (`prev`).
void engine_out(const Tf64State &s, vsmcx_threefry4x64_state &e,
                const vsmcx_threefry4x64_state *prev = nullptr)
{
    bool drew = true;
    if (prev) {
        drew = s.index != (prev->index > 4 ? 4u : (uint32_t)prev->index);
        for (int j = 0; j < 4; ++j) drew = drew || s.ctr[j] != prev->ctr[j];
    }
    uint64_t buf[4] = {0, 0, 0, 0};
    if (!drew) {
        for (int j = 0; j < 4; ++j) buf[j] = prev->buffer[j];
    } else if (prev || s.index < 4) {
        u64x4 c;
        for (int j = 0; j < 4; ++j) c.v[j] = s.ctr[j];
        const u64x4 b = threefry4x64_20(c, s.key);
        for (int j = 0; j < 4; ++j) buf[j] = b.v[j];
    }
    for (int j = 0; j < 4; ++j) {
        e.ctr[j] = s.ctr[j];
        e.key[j] = s.key[j];
        e.buffer[j] = buf[j];
    }
    e.index = s.index;
}


}  // namespace

namespace vsmcx_host {

bool resample_args(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                   const vsmcx_threefry4x64_state *eng, const double *weight,
                   const uint64_t *replication)
{
    REQUIRE(ctx && eng, "NULL argument");
    REQUIRE(scheme_ok(scheme), "unknown resampling scheme");
    if (M == 0) return false;
    REQUIRE(weight && replication, "NULL weight/replication");
    REQUIRE(M < (1ull << 31) && N < (1ull << 31), "sizes must be below 2^31");
    for (uint64_t i = 0; i < M; ++i)
        REQUIRE(weight[i] >= 0 && std::isfinite(weight[i]), "weights must be finite and >= 0");
    return true;
}

void resample_on(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                 vsmcx_threefry4x64_state *eng, const double *weight, uint64_t *replication,
                 uint64_t *copy_from, ResampleBuffers &rb, FilterCtl *ctl, double *w,
                 uint64_t *cf_dev)
{
    const bool lists = copy_from && M == N;
    FilterCtl h{};
    h.n = N;
    h.resample = 1;
    h.scheme = scheme;
    h.eng = engine_in(*eng);
    VSMCX_HIP(hipMemcpyAsync(ctl, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    VSMCX_HIP(hipMemcpyAsync(w, weight, 8 * M, hipMemcpyHostToDevice, ctx->stream));
    launch_resample_from_weights(ctl, w, M, scheme, rb, nullptr, ctx->stream);
    VSMCX_HIP(hipGetLastError());
    VSMCX_HIP(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    VSMCX_HIP(hipMemcpyAsync(replication, rb.rep_out, 8 * M, hipMemcpyDeviceToHost,
                             ctx->stream));
    std::vector<uint32_t> anc;
    if (lists && cf_dev) {
        // widened on the device, copied straight into the caller's array
        launch_widen_u32(rb.anc, cf_dev, M, ctx->stream);
        VSMCX_HIP(hipGetLastError());
        VSMCX_HIP(hipMemcpyAsync(copy_from, cf_dev, 8 * M, hipMemcpyDeviceToHost, ctx->stream));
    } else if (lists) {
        anc.resize(M);
        VSMCX_HIP(hipMemcpyAsync(anc.data(), rb.anc, 4 * M, hipMemcpyDeviceToHost, ctx->stream));
    }
    ctx->sync();
    check_dev_err(h);
    if (copy_from) {
        if (lists) {
            if (!cf_dev)
                for (uint64_t i = 0; i < M; ++i) copy_from[i] = anc[i];
        } else {  // cfrp_trans M != N branch: expansion (common.hpp:110-117)
            uint64_t to = 0;
            for (uint64_t from = 0; from < M; ++from)
                for (uint64_t r = 0; r < replication[from] && to < N; ++r)
                    copy_from[to++] = from;
        }
    }
    const vsmcx_threefry4x64_state in = *eng;
    engine_out(h.eng, *eng, &in);
}

}  // namespace vsmcx_host

struct vsmcx_filter {
    vsmcx_ctx *ctx = nullptr;
    vsmcx_filter_desc d{};
    CvParams p{};
    bool fp64 = false;
    uint64_t N = 0, T = 0, tiles = 0, nobs = 0, t = 0;
    void *st[2] = {nullptr, nullptr};
    int cur = 0;
    double *lw = nullptr;
    float *lwc = nullptr;  // compact log-weights (mixed pipeline)
    double *partials = nullptr;
    double *obs = nullptr;
    FilterCtl *ctl = nullptr;
    StepRow *hist = nullptr;
    ResampleBuffers rb;
    std::vector<hipEvent_t> events;
    // move-kernel profiling: event pairs around each move launch
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pool;
    size_t prof_used = 0;
    // plain FP64 pipeline only
    bool plain = false;
    double *w = nullptr, *inc = nullptr, *ppart = nullptr, *scal = nullptr;

    size_t state_bytes() const { return fp64 ? 32 : 16; }

    // the move kernel of step t (the dominant kernel of the pipeline); a
    // multinomial filter forks its step's sorted uniforms to the side stream
    // first, so they overlap the move and the weight fold's first passes
    void launch_move()
    {
        hipStream_t s = ctx->stream;
        if (rb.side.s && d.scheme == kMultinomial) launch_side_uniforms(ctl, rb, s);
        if (plain)
            launch_plain_move(st[0], inc, obs, p, d.seed, t, N, s);
        else
            launch_cv_move(fp64, ctl, st[cur], st[1 - cur], rb.anc, rb.sl, lw, obs, p, d.seed, t, N,
                           partials, s);
    }
    // everything after the move: weights, decision, resampling, monitor
    void launch_rest()
    {
        hipStream_t s = ctx->stream;
        if (plain) {
            launch_plain_rest(ctl, st[0], lw, w, inc, ppart, scal, t, N, d.scheme, rb, hist, s);
        } else {
            // non-residual schemes: the reduce also runs the resample prologue and
            // the fold's tile starts
            const bool fuse = d.scheme == kMultinomial || d.scheme == kStratified ||
                              d.scheme == kSystematic;
            if (fuse) {
                rb.fs.tiles = tiles;
                launch_cv_reduce(ctl, partials, tiles, hist, t, s, d.scheme, &rb.fs);
            } else {
                launch_cv_reduce(ctl, partials, tiles, hist, t, s);
            }
            // encoded ancestors: the next move resolves zero slots through rb.sl
            launch_resample_from_log_weights(ctl, lw, N, d.scheme, rb, hist, s, partials, fuse,
                                             true);
            cur ^= 1;
        }
        ++t;
    }
    void step()
    {
        if (t >= nobs) throw DevError(VSMCX_ERR_STATE, "no observation for step " + std::to_string(t));
        if (profiling) {
            if (prof_used == prof_pool.size()) {
                hipEvent_t a, b;
                VSMCX_HIP(hipEventCreate(&a));
                VSMCX_HIP(hipEventCreate(&b));
                prof_pool.emplace_back(a, b);
            }
            auto &ev = prof_pool[prof_used++];
            VSMCX_HIP(hipEventRecord(ev.first, ctx->stream));
            launch_move();
            VSMCX_HIP(hipEventRecord(ev.second, ctx->stream));
        } else {
            launch_move();
        }
        launch_rest();
    }
};

struct vsmcx_timer {
    vsmcx_ctx *ctx = nullptr;
    std::vector<hipEvent_t> ev;
};

extern "C" {

const char *vsmcx_last_error(void) { return g_err.c_str(); }
const char *vsmcx_version(void) { return "vsmcx 0.1 (gfx950)"; }

int vsmcx_ctx_create(int device, uint64_t hbm_budget, vsmcx_ctx **out)
{
    API_BEGIN
    REQUIRE(out, "out is NULL");
    int n = 0;
    VSMCX_HIP(hipGetDeviceCount(&n));
    REQUIRE(device >= 0 && device < n, "no such device");
    auto c = std::make_unique<vsmcx_ctx>();
    c->device = device;
    c->budget = hbm_budget;
    VSMCX_HIP(hipSetDevice(device));
    c->scratch = noalloc_scratch(device);
    VSMCX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c.release();
    API_END
}

int vsmcx_ctx_destroy(vsmcx_ctx *ctx)
{
    API_BEGIN
    if (!ctx) return VSMCX_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->sizes) (void)hipFree(kv.first);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    API_END
}

int vsmcx_ctx_sync(vsmcx_ctx *ctx)
{
    API_BEGIN
    REQUIRE(ctx, "ctx is NULL");
    ctx->sync();
    API_END
}

int vsmcx_ctx_allocated(vsmcx_ctx *ctx, uint64_t *bytes)
{
    API_BEGIN
    REQUIRE(ctx && bytes, "NULL argument");
    *bytes = ctx->allocated;
    API_END
}

int vsmcx_stream_probe(vsmcx_ctx *ctx, uint64_t bytes, int reps, double gbps[3])
{
    API_BEGIN
    REQUIRE(ctx && gbps, "NULL argument");
    REQUIRE(bytes >= (1ull << 20) && bytes <= (1ull << 36), "bytes outside [2^20, 2^36]");
    REQUIRE(reps >= 1 && reps <= 100, "reps outside [1, 100]");
    const uint64_t n = bytes / 16;
    float4 *a = ctx->alloc_n<float4>(n);
    float4 *b = nullptr;
    try {
        b = ctx->alloc_n<float4>(n);
    } catch (...) {
        ctx->release(a);
        throw;
    }
    hipError_t e = hipMemsetAsync(a, 0, n * 16, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b, 0, n * 16, ctx->stream);
    float ms[3] = {-1.f, -1.f, -1.f};
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        ms[k] = launch_stream_probe(k, a, b, n, reps, ctx->stream);
        if (ms[k] <= 0.f) e = hipErrorUnknown;
    }
    // nothing may still use the buffers when they are released, error or not
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    ctx->release(a);
    ctx->release(b);
    VSMCX_HIP(e);
    const double nb = (double)(n * 16);
    gbps[0] = 2 * nb / (ms[0] * 1e-3) / 1e9;
    gbps[1] = nb / (ms[1] * 1e-3) / 1e9;
    gbps[2] = nb / (ms[2] * 1e-3) / 1e9;
    API_END
}

// ---------------------------------------------------------------- RNG
int vsmcx_philox4x32_fill(vsmcx_ctx *ctx, const uint32_t key[2], const uint32_t ctr0[4],
                          uint64_t nblocks, uint32_t *out)
{
    API_BEGIN
    REQUIRE(ctx && key && ctr0 && (out || !nblocks), "NULL argument");
    if (!nblocks) return VSMCX_OK;
    uint32_t *d = ctx->alloc_n<uint32_t>(4 * nblocks);
    launch_philox4x32_fill(key[0], key[1], ctr0, nblocks, d, ctx->stream);
    VSMCX_HIP(hipMemcpyAsync(out, d, 16 * nblocks, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    ctx->release(d);
    API_END
}

int vsmcx_threefry4x64_fill(vsmcx_ctx *ctx, const uint64_t key[4], const uint64_t ctr0[4],
                            uint64_t nblocks, uint64_t *out)
{
    API_BEGIN
    REQUIRE(ctx && key && ctr0 && (out || !nblocks), "NULL argument");
    if (!nblocks) return VSMCX_OK;
    uint64_t *d = ctx->alloc_n<uint64_t>(4 * nblocks);
    launch_threefry4x64_fill(key, ctr0, nblocks, d, ctx->stream);
    VSMCX_HIP(hipMemcpyAsync(out, d, 32 * nblocks, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    ctx->release(d);
    API_END
}

int vsmcx_threefry4x32_fill(vsmcx_ctx *ctx, const uint32_t key[4], const uint32_t ctr0[4],
                            uint64_t nblocks, uint32_t *out)
{
    API_BEGIN
    REQUIRE(ctx && key && ctr0 && (out || !nblocks), "NULL argument");
    if (!nblocks) return VSMCX_OK;
    uint32_t *d = ctx->alloc_n<uint32_t>(4 * nblocks);
    launch_threefry4x32_fill(key, ctr0, nblocks, d, ctx->stream);
    VSMCX_HIP(hipMemcpyAsync(out, d, 16 * nblocks, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    ctx->release(d);
    API_END
}

// ---------------------------------------------------------------- resample
int vsmcx_resample(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                   vsmcx_threefry4x64_state *eng, const double *weight, uint64_t *replication,
                   uint64_t *copy_from)
{
    API_BEGIN
    if (!resample_args(ctx, scheme, M, N, eng, weight, replication)) return VSMCX_OK;
    const bool lists = copy_from && M == N;
    ResampleBuffers rb;
    FilterCtl *ctl = nullptr;
    double *w = nullptr;
    try {
        alloc_resample(ctx, rb, M, N, lists);
        rb.rep_out = ctx->alloc_n<uint64_t>(M);
        ctl = ctx->alloc_n<FilterCtl>(1);
        w = ctx->alloc_n<double>(M);
        resample_on(ctx, scheme, M, N, eng, weight, replication, copy_from, rb, ctl, w, nullptr);
    } catch (...) {
        free_resample(ctx, rb);
        ctx->release(ctl);
        ctx->release(w);
        throw;
    }
    free_resample(ctx, rb);
    ctx->release(ctl);
    ctx->release(w);
    API_END
}

int vsmcx_fold_prefix(vsmcx_ctx *ctx, uint64_t n, const double *x, double *out)
{
    API_BEGIN
    REQUIRE(ctx, "NULL ctx");
    if (n == 0) return VSMCX_OK;
    REQUIRE(x && out, "NULL argument");
    for (uint64_t i = 0; i < n; ++i)
        REQUIRE(x[i] >= 0 && std::isfinite(x[i]), "fold terms must be finite and >= 0");
    ResampleBuffers rb;
    FilterCtl *ctl = nullptr;
    double *dx = nullptr;
    try {
        alloc_resample(ctx, rb, n, 1, false);
        rb.a_out = ctx->alloc_n<double>(n);
        ctl = ctx->alloc_n<FilterCtl>(1);
        dx = ctx->alloc_n<double>(n);
        FilterCtl h{};
        h.n = n;
        h.resample = 1;
        VSMCX_HIP(hipMemcpyAsync(ctl, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
        VSMCX_HIP(hipMemcpyAsync(dx, x, 8 * n, hipMemcpyHostToDevice, ctx->stream));
        launch_fold_values(ctl, dx, n, rb, ctx->stream);
        VSMCX_HIP(hipGetLastError());
        VSMCX_HIP(hipMemcpyAsync(out, rb.a_out, 8 * n, hipMemcpyDeviceToHost, ctx->stream));
        VSMCX_HIP(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        check_dev_err(h);
    } catch (...) {
        free_resample(ctx, rb);
        ctx->release(ctl);
        ctx->release(dx);
        throw;
    }
    free_resample(ctx, rb);
    ctx->release(ctl);
    ctx->release(dx);
    API_END
}

// ---------------------------------------------------------------- filter
int vsmcx_filter_create(vsmcx_ctx *ctx, const vsmcx_filter_desc *desc, vsmcx_filter **out)
{
    API_BEGIN
    REQUIRE(ctx && desc && out, "NULL argument");
    REQUIRE(desc->n >= 1 && desc->n < (1ull << 31), "n must be in [1, 2^31)");
    REQUIRE(desc->horizon >= 1, "horizon must be >= 1");
    REQUIRE(scheme_ok(desc->scheme), "unknown resampling scheme");
    REQUIRE(desc->precision == VSMCX_MIXED || desc->precision == VSMCX_FP64_FUSED ||
                desc->precision == VSMCX_FP64_PLAIN,
            "precision must be VSMCX_MIXED, VSMCX_FP64_FUSED or VSMCX_FP64_PLAIN");
    auto f = std::make_unique<vsmcx_filter>();
    f->ctx = ctx;
    f->d = *desc;
    f->fp64 = desc->precision != VSMCX_MIXED;
    f->plain = desc->precision == VSMCX_FP64_PLAIN;
    f->p = CvParams{desc->dt, desc->sd_pos, desc->sd_vel, desc->prior_sd_pos,
                    desc->prior_sd_vel, desc->obs_scale, desc->nu, desc->llh_const};
    f->N = desc->n;
    f->T = desc->horizon;
    f->tiles = (f->N + kMoveTile - 1) / kMoveTile;
    try {
        f->st[0] = ctx->alloc(f->N * f->state_bytes());
        if (!f->plain) f->st[1] = ctx->alloc(f->N * f->state_bytes());
        if (f->plain) {
            f->w = ctx->alloc_n<double>(f->N);
            f->inc = ctx->alloc_n<double>(f->N);
            f->ppart = ctx->alloc_n<double>(plain_partials(f->N));
            f->scal = ctx->alloc_zero<double>(8);
            launch_plain_init_weights(f->w, f->N, ctx->stream);
        }
        f->lw = ctx->alloc_n<double>(f->N);
        if (!f->fp64 && !f->plain) f->lwc = ctx->alloc_n<float>(f->N);
        f->partials = ctx->alloc_n<double>(5 * f->tiles);
        f->obs = ctx->alloc_n<double>(2 * f->T);
        f->ctl = ctx->alloc_n<FilterCtl>(1);
        f->hist = ctx->alloc_zero<StepRow>(f->T);
        // the fused pipelines run the sorted uniforms' fold on a side stream
        const bool side = !f->plain && (desc->scheme == VSMCX_MULTINOMIAL ||
                                        desc->scheme == VSMCX_RESIDUAL);
        alloc_resample(ctx, f->rb, f->N, f->N, true, side);
        if (side) {
            SideFold &sd = f->rb.side;
            sd.uctl = ctx->alloc_zero<FilterCtl>(1);
            VSMCX_HIP(hipStreamCreateWithFlags(&sd.s, hipStreamNonBlocking));
            VSMCX_HIP(hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming));
            VSMCX_HIP(hipEventCreateWithFlags(&sd.join, hipEventDisableTiming));
            VSMCX_HIP(hipEventCreateWithFlags(&sd.starts, hipEventDisableTiming));
        }
    } catch (...) {
        for (void *p : {f->st[0], f->st[1], (void *)f->lw, (void *)f->lwc, (void *)f->partials, (void *)f->obs,
                        (void *)f->ctl, (void *)f->hist, (void *)f->w, (void *)f->inc,
                        (void *)f->ppart, (void *)f->scal})
            ctx->release(p);
        free_side(ctx, f->rb.side);
        free_resample(ctx, f->rb);
        throw;
    }
    FilterCtl h{};
    h.n = f->N;
    h.threshold = desc->threshold;
    h.scheme = desc->scheme;
    h.eng.key[0] = desc->resample_seed;
    h.eng.index = 4;
    h.lwc = f->lwc;
    VSMCX_HIP(hipMemcpyAsync(f->ctl, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
    *out = f.release();
    API_END
}

int vsmcx_filter_destroy(vsmcx_filter *f)
{
    API_BEGIN
    if (!f) return VSMCX_OK;
    vsmcx_ctx *c = f->ctx;
    (void)hipStreamSynchronize(c->stream);
    free_side(c, f->rb.side);
    for (void *p : {f->st[0], f->st[1], (void *)f->lw, (void *)f->lwc, (void *)f->partials, (void *)f->obs,
                    (void *)f->ctl, (void *)f->hist, (void *)f->w, (void *)f->inc,
                    (void *)f->ppart, (void *)f->scal})
        c->release(p);
    free_resample(c, f->rb);
    for (hipEvent_t e : f->events) (void)hipEventDestroy(e);
    for (auto &pr : f->prof_pool) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    delete f;
    API_END
}

int vsmcx_filter_set_obs(vsmcx_filter *f, uint64_t T, const double *obs)
{
    API_BEGIN
    REQUIRE(f && obs, "NULL argument");
    REQUIRE(T <= f->T, "more observations than the horizon");
    VSMCX_HIP(hipMemcpyAsync(f->obs, obs, 16 * T, hipMemcpyHostToDevice, f->ctx->stream));
    f->nobs = T;
    f->ctx->sync();
    API_END
}

int vsmcx_filter_run(vsmcx_filter *f, uint64_t nsteps)
{
    API_BEGIN
    REQUIRE(f, "NULL filter");
    REQUIRE(f->t + nsteps <= f->nobs, "not enough observations for the requested steps");
    for (uint64_t k = 0; k < nsteps; ++k) f->step();
    VSMCX_HIP(hipGetLastError());
    API_END
}

int vsmcx_filter_finish(vsmcx_filter *f)
{
    API_BEGIN
    REQUIRE(f, "NULL filter");
    if (f->t > 0 && !f->plain)
        launch_cv_monitor(f->fp64, f->ctl, f->st[f->cur], f->rb.anc, f->rb.sl, f->lw, f->N,
                          f->partials,
                          f->hist, f->ctx->stream);
    VSMCX_HIP(hipGetLastError());
    FilterCtl h;
    VSMCX_HIP(hipMemcpyAsync(&h, f->ctl, sizeof(h), hipMemcpyDeviceToHost, f->ctx->stream));
    f->ctx->sync();
    check_dev_err(h);
    API_END
}

int vsmcx_filter_steps(vsmcx_filter *f, uint64_t *t)
{
    API_BEGIN
    REQUIRE(f && t, "NULL argument");
    *t = f->t;
    API_END
}

int vsmcx_filter_history(vsmcx_filter *f, uint64_t count, vsmcx_step_stats *out)
{
    API_BEGIN
    static_assert(sizeof(StepRow) == sizeof(vsmcx_step_stats), "history row layout");
    REQUIRE(f && out, "NULL argument");
    REQUIRE(count <= f->t, "count beyond completed steps");
    VSMCX_HIP(hipMemcpyAsync(out, f->hist, sizeof(StepRow) * count, hipMemcpyDeviceToHost,
                             f->ctx->stream));
    f->ctx->sync();
    API_END
}

// ---- checkpoint / resume --------------------------------------------------------
// A fused filter's whole device state in one file of this build's own layout:
//   "VSMCXCK1", u32 version, u32 sizeof(FilterCtl), u32 sizeof(StepRow),
//   u32 state bytes per particle, vsmcx_filter_desc, u64 t, u64 nobs,
//   FilterCtl (weights' scalars, log-evidence, the resample engine, flags),
//   state rows (N), log-weights (N f64), ancestors (N u32),
//   observations (nobs x 2 f64), history (horizon rows).
// Loading creates the filter and resumes it where the save left it: the
// steps after the load compute exactly what an uninterrupted run computes.
extern "C++" {
namespace {
const char kCkMagic[8] = {'V', 'S', 'M', 'C', 'X', 'C', 'K', '1'};
const uint32_t kCkVersion = 1;

struct CkHeader {
    char magic[8];
    uint32_t version, ctl_bytes, row_bytes, state_bytes;
    vsmcx_filter_desc desc;
    uint64_t t, nobs;
};

struct File {
    std::FILE *fp;
    File(const char *path, const char *mode) : fp(std::fopen(path, mode))
    {
        if (!fp) throw ArgError(std::string("cannot open ") + path);
    }
    ~File()
    {
        if (fp) std::fclose(fp);
    }
    void put(const void *p, size_t n)
    {
        if (n && std::fwrite(p, 1, n, fp) != n) throw DevError(VSMCX_ERR_INTERNAL, "checkpoint write failed");
    }
    void get(void *p, size_t n)
    {
        if (n && std::fread(p, 1, n, fp) != n) throw ArgError("checkpoint file is truncated");
    }
    void close()
    {
        const int rc = std::fclose(fp);
        fp = nullptr;
        if (rc != 0) throw DevError(VSMCX_ERR_INTERNAL, "checkpoint write failed");
    }
};

template <typename T>
std::vector<T> to_host(vsmcx_ctx *c, const void *d, size_t n)
{
    std::vector<T> h(n);
    if (n) VSMCX_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    return h;
}
}  // namespace
}  // extern "C++"

int vsmcx_filter_describe(vsmcx_filter *f, vsmcx_filter_desc *out)
{
    API_BEGIN
    REQUIRE(f && out, "NULL argument");
    *out = f->d;
    API_END
}

int vsmcx_filter_save(vsmcx_filter *f, const char *path)
{
    API_BEGIN
    REQUIRE(f && path, "NULL argument");
    REQUIRE(!f->plain, "checkpoints need a fused pipeline (VSMCX_MIXED or VSMCX_FP64_FUSED)");
    const uint64_t N = f->N;
    CkHeader h{};
    std::memcpy(h.magic, kCkMagic, 8);
    h.version = kCkVersion;
    h.ctl_bytes = sizeof(FilterCtl);
    h.row_bytes = sizeof(StepRow);
    h.state_bytes = (uint32_t)f->state_bytes();
    h.desc = f->d;
    h.t = f->t;
    h.nobs = f->nobs;
    FilterCtl ctl;
    VSMCX_HIP(hipMemcpyAsync(&ctl, f->ctl, sizeof(ctl), hipMemcpyDeviceToHost, f->ctx->stream));
    auto st = to_host<unsigned char>(f->ctx, f->st[f->cur], N * f->state_bytes());
    // the file holds FP64 log-weights whatever their form on the device
    launch_lw_expand(f->ctl, f->lw, N, f->rb.fs.wx, f->ctx->stream);
    auto lw = to_host<double>(f->ctx, f->rb.fs.wx, N);
    auto anc = to_host<uint32_t>(f->ctx, f->rb.anc, N);
    auto sl = to_host<uint32_t>(f->ctx, f->rb.sl, N);
    auto obs = to_host<double>(f->ctx, f->obs, 2 * f->nobs);
    auto hist = to_host<StepRow>(f->ctx, f->hist, f->T);
    f->ctx->sync();
    check_dev_err(ctl);
    ctl.lw_kind = 0;  // the file holds FP64 log-weights and no device pointer
    ctl.lw_store_kind = 0;
    ctl.lwc = nullptr;
    // the file holds decoded ancestors (every slot's source)
    if (ctl.prev_resampled)
        for (uint64_t i = 0; i < N; ++i) anc[i] = anc_source(anc.data(), sl.data(), i);
    // written beside the target and renamed over it, so a failed write
    // leaves an earlier checkpoint at `path` intact
    const std::string tmp = std::string(path) + ".tmp";
    try {
        File out(tmp.c_str(), "wb");
        out.put(&h, sizeof(h));
        out.put(&ctl, sizeof(ctl));
        out.put(st.data(), st.size());
        out.put(lw.data(), 8 * N);
        out.put(anc.data(), 4 * N);
        out.put(obs.data(), 8 * obs.size());
        out.put(hist.data(), sizeof(StepRow) * hist.size());
        out.close();
    } catch (...) {
        std::remove(tmp.c_str());
        throw;
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        throw DevError(VSMCX_ERR_INTERNAL, std::string("cannot rename the checkpoint to ") + path);
    }
    API_END
}

int vsmcx_filter_load(vsmcx_ctx *ctx, const char *path, vsmcx_filter **out)
{
    API_BEGIN
    REQUIRE(ctx && path && out, "NULL argument");
    File in(path, "rb");
    CkHeader h;
    in.get(&h, sizeof(h));
    REQUIRE(std::memcmp(h.magic, kCkMagic, 8) == 0, "not a vsmcx filter checkpoint");
    REQUIRE(h.version == kCkVersion && h.ctl_bytes == sizeof(FilterCtl) &&
                h.row_bytes == sizeof(StepRow),
            "checkpoint written by an incompatible build");
    REQUIRE(h.desc.precision == VSMCX_MIXED || h.desc.precision == VSMCX_FP64_FUSED,
            "checkpoint of an unsupported pipeline");
    REQUIRE(h.state_bytes == (h.desc.precision == VSMCX_MIXED ? 16u : 32u),
            "checkpoint state size does not match its precision");
    REQUIRE(h.nobs <= h.desc.horizon && h.t <= h.nobs, "checkpoint counters out of range");
    // sizes checked against the file before anything is allocated from them
    REQUIRE(h.desc.n >= 1 && h.desc.n < (1ull << 31), "checkpoint particle count out of range");
    REQUIRE(h.desc.horizon >= 1 && h.desc.horizon < (1ull << 32),
            "checkpoint horizon out of range");
    REQUIRE(scheme_ok(h.desc.scheme), "checkpoint of an unknown resampling scheme");
    const uint64_t N = h.desc.n;
    const uint64_t payload = sizeof(CkHeader) + sizeof(FilterCtl) + N * h.state_bytes + 12 * N +
                             16 * h.nobs + sizeof(StepRow) * h.desc.horizon;
    REQUIRE(std::fseek(in.fp, 0, SEEK_END) == 0, "cannot read the checkpoint");
    const long fsize = std::ftell(in.fp);
    REQUIRE(fsize >= 0 && (uint64_t)fsize == payload, "checkpoint file size does not match its header");
    REQUIRE(std::fseek(in.fp, (long)sizeof(CkHeader), SEEK_SET) == 0, "cannot read the checkpoint");
    FilterCtl ctl;
    in.get(&ctl, sizeof(ctl));
    REQUIRE(ctl.n == N, "checkpoint control block does not match its header");
    std::vector<unsigned char> st(N * h.state_bytes);
    std::vector<double> lw(N), obs(2 * h.nobs);
    std::vector<uint32_t> anc(N);
    std::vector<StepRow> hist(h.desc.horizon);
    in.get(st.data(), st.size());
    in.get(lw.data(), 8 * N);
    in.get(anc.data(), 4 * N);
    in.get(obs.data(), 8 * obs.size());
    in.get(hist.data(), sizeof(StepRow) * hist.size());
    vsmcx_filter *f = nullptr;
    const int rc = vsmcx_filter_create(ctx, &h.desc, &f);
    if (rc != VSMCX_OK) return rc;
    std::unique_ptr<vsmcx_filter, int (*)(vsmcx_filter *)> guard(f, vsmcx_filter_destroy);
    hipStream_t s = ctx->stream;
    VSMCX_HIP(hipMemcpyAsync(f->st[0], st.data(), st.size(), hipMemcpyHostToDevice, s));
    VSMCX_HIP(hipMemcpyAsync(f->lw, lw.data(), 8 * N, hipMemcpyHostToDevice, s));
    VSMCX_HIP(hipMemcpyAsync(f->rb.anc, anc.data(), 4 * N, hipMemcpyHostToDevice, s));
    ctl.lw_kind = 0;
    ctl.lw_store_kind = 0;
    ctl.lwc = f->lwc;
    if (h.nobs)
        VSMCX_HIP(hipMemcpyAsync(f->obs, obs.data(), 8 * obs.size(), hipMemcpyHostToDevice, s));
    VSMCX_HIP(hipMemcpyAsync(f->hist, hist.data(), sizeof(StepRow) * hist.size(),
                             hipMemcpyHostToDevice, s));
    VSMCX_HIP(hipMemcpyAsync(f->ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, s));
    ctx->sync();
    f->cur = 0;
    f->t = h.t;
    f->nobs = h.nobs;
    *out = guard.release();
    API_END
}

int vsmcx_filter_state(vsmcx_filter *f, double *state, double *log_weight)
{
    API_BEGIN
    REQUIRE(f, "NULL filter");
    const uint64_t N = f->N;
    if (state && f->plain) {
        VSMCX_HIP(hipMemcpyAsync(state, f->st[0], 32 * N, hipMemcpyDeviceToHost, f->ctx->stream));
        f->ctx->sync();
    } else if (state) {
        void *tmp = f->ctx->alloc(N * f->state_bytes());
        launch_cv_gather(f->fp64, f->ctl, f->st[f->cur], tmp, f->rb.anc, f->rb.sl, N,
                         f->ctx->stream);
        if (f->fp64) {
            VSMCX_HIP(hipMemcpyAsync(state, tmp, 32 * N, hipMemcpyDeviceToHost, f->ctx->stream));
            f->ctx->sync();
        } else {
            std::vector<float> h(4 * N);
            VSMCX_HIP(hipMemcpyAsync(h.data(), tmp, 16 * N, hipMemcpyDeviceToHost, f->ctx->stream));
            f->ctx->sync();
            for (uint64_t i = 0; i < 4 * N; ++i) state[i] = h[i];
        }
        f->ctx->release(tmp);
    }
    if (log_weight) {
        FilterCtl h;
        VSMCX_HIP(hipMemcpyAsync(&h, f->ctl, sizeof(h), hipMemcpyDeviceToHost, f->ctx->stream));
        launch_lw_expand(f->ctl, f->lw, N, f->rb.fs.wx, f->ctx->stream);
        VSMCX_HIP(hipMemcpyAsync(log_weight, f->rb.fs.wx, 8 * N, hipMemcpyDeviceToHost,
                                 f->ctx->stream));
        f->ctx->sync();
        // vSMC keeps log-weights normalised to max 0 (core/weight_set.hpp:108-117);
        // after a resample they are all zero (set_equal_weight).  The plain
        // pipeline stores them that way already.
        if (!f->plain)
            for (uint64_t i = 0; i < N; ++i)
                log_weight[i] = h.prev_resampled ? 0.0 : log_weight[i] - h.m;
    }
    API_END
}

int vsmcx_filter_engine(vsmcx_filter *f, vsmcx_threefry4x64_state *eng)
{
    API_BEGIN
    REQUIRE(f && eng, "NULL argument");
    FilterCtl h;
    VSMCX_HIP(hipMemcpyAsync(&h, f->ctl, sizeof(h), hipMemcpyDeviceToHost, f->ctx->stream));
    f->ctx->sync();
    vsmcx_threefry4x64_state seeded{};  // the engine as the filter seeded it
    seeded.key[0] = f->d.resample_seed;
    seeded.index = 4;
    engine_out(h.eng, *eng, &seeded);
    API_END
}

int vsmcx_filter_time(vsmcx_filter *f, uint64_t nsteps, double *elapsed_ms, double *move_kernel_ms)
{
    API_BEGIN
    REQUIRE(f && elapsed_ms, "NULL argument");
    REQUIRE(f->t + nsteps <= f->nobs, "not enough observations for the requested steps");
    const size_t need = 2 * nsteps + 2;
    while (f->events.size() < need) {
        hipEvent_t e;
        VSMCX_HIP(hipEventCreate(&e));
        f->events.push_back(e);
    }
    hipStream_t s = f->ctx->stream;
    VSMCX_HIP(hipEventRecord(f->events[0], s));
    for (uint64_t k = 0; k < nsteps; ++k) {
        if (f->t >= f->nobs) throw DevError(VSMCX_ERR_STATE, "out of observations");
        f->launch_move();
        VSMCX_HIP(hipEventRecord(f->events[2 + 2 * k], s));
        f->launch_rest();
        VSMCX_HIP(hipEventRecord(f->events[3 + 2 * k], s));
    }
    VSMCX_HIP(hipEventRecord(f->events[1], s));
    VSMCX_HIP(hipGetLastError());
    VSMCX_HIP(hipEventSynchronize(f->events[1]));
    float ms = 0;
    VSMCX_HIP(hipEventElapsedTime(&ms, f->events[0], f->events[1]));
    *elapsed_ms = ms;
    if (move_kernel_ms) {
        // move k runs between the end event of step k-1 (or the start) and events[2+2k]
        double tot = 0;
        for (uint64_t k = 0; k < nsteps; ++k) {
            float a = 0;
            hipEvent_t prev = k == 0 ? f->events[0] : f->events[3 + 2 * (k - 1)];
            VSMCX_HIP(hipEventElapsedTime(&a, prev, f->events[2 + 2 * k]));
            tot += a;
        }
        *move_kernel_ms = nsteps ? tot / nsteps : 0;
    }
    FilterCtl h;
    VSMCX_HIP(hipMemcpy(&h, f->ctl, sizeof(h), hipMemcpyDeviceToHost));
    check_dev_err(h);
    API_END
}

int vsmcx_timer_create(vsmcx_ctx *ctx, uint64_t n, vsmcx_timer **out)
{
    API_BEGIN
    REQUIRE(ctx && out && n >= 1, "bad argument");
    auto tm = std::make_unique<vsmcx_timer>();
    tm->ctx = ctx;
    tm->ev.resize(n);
    for (auto &e : tm->ev) VSMCX_HIP(hipEventCreate(&e));
    *out = tm.release();
    API_END
}

int vsmcx_timer_record(vsmcx_timer *tm, uint64_t i)
{
    API_BEGIN
    REQUIRE(tm && i < tm->ev.size(), "bad timer mark");
    VSMCX_HIP(hipEventRecord(tm->ev[i], tm->ctx->stream));
    API_END
}

int vsmcx_timer_elapsed(vsmcx_timer *tm, uint64_t i, uint64_t j, double *ms)
{
    API_BEGIN
    REQUIRE(tm && ms && i < tm->ev.size() && j < tm->ev.size(), "bad timer mark");
    VSMCX_HIP(hipEventSynchronize(tm->ev[j]));
    float v = 0;
    VSMCX_HIP(hipEventElapsedTime(&v, tm->ev[i], tm->ev[j]));
    *ms = v;
    API_END
}

int vsmcx_timer_destroy(vsmcx_timer *tm)
{
    API_BEGIN
    if (!tm) return VSMCX_OK;
    for (auto &e : tm->ev) (void)hipEventDestroy(e);
    delete tm;
    API_END
}

int vsmcx_filter_profile(vsmcx_filter *f, int enable)
{
    API_BEGIN
    REQUIRE(f, "NULL filter");
    f->profiling = enable != 0;
    f->prof_used = 0;
    API_END
}

int vsmcx_filter_move_times(vsmcx_filter *f, uint64_t max, uint64_t *count, double *ms)
{
    API_BEGIN
    REQUIRE(f && count, "NULL argument");
    f->ctx->sync();
    const uint64_t n = f->prof_used < max ? f->prof_used : max;
    for (uint64_t k = 0; k < n; ++k) {
        float v = 0;
        VSMCX_HIP(hipEventElapsedTime(&v, f->prof_pool[k].first, f->prof_pool[k].second));
        if (ms) ms[k] = v;
    }
    *count = n;
    f->prof_used = 0;
    API_END
}

int vsmcx_filter_debug_step(vsmcx_filter *f, double *weight, uint64_t *replication,
                            uint64_t *copy_from, double *uniforms, uint64_t *nn,
                            vsmcx_threefry4x64_state *eng_before, int *resampled)
{
    API_BEGIN
    REQUIRE(f, "NULL filter");
    REQUIRE(!f->plain, "debug_step needs a fused pipeline");
    REQUIRE(f->t < f->nobs, "out of observations");
    vsmcx_ctx *c = f->ctx;
    hipStream_t s = c->stream;
    const uint64_t N = f->N;
    launch_cv_move(f->fp64, f->ctl, f->st[f->cur], f->st[1 - f->cur], f->rb.anc, f->rb.sl, f->lw,
                   f->obs,
                   f->p, f->d.seed, f->t, N, f->partials, s);
    launch_cv_reduce(f->ctl, f->partials, f->tiles, f->hist, f->t, s);
    FilterCtl h;
    VSMCX_HIP(hipMemcpyAsync(&h, f->ctl, sizeof(h), hipMemcpyDeviceToHost, s));
    double *w = c->alloc_n<double>(N);
    launch_cv_weights(f->ctl, f->lw, N, w, s);
    if (weight) VSMCX_HIP(hipMemcpyAsync(weight, w, 8 * N, hipMemcpyDeviceToHost, s));
    c->sync();
    c->release(w);
    if (eng_before) {
        vsmcx_threefry4x64_state seeded{};
        seeded.key[0] = f->d.resample_seed;
        seeded.index = 4;
        engine_out(h.eng, *eng_before, &seeded);
    }
    if (resampled) *resampled = h.resample;
    f->rb.rep_out = c->alloc_n<uint64_t>(N);
    launch_resample_from_log_weights(f->ctl, f->lw, N, f->d.scheme, f->rb, f->hist, s, f->partials);
    VSMCX_HIP(hipGetLastError());
    if (h.resample) {
        if (replication)
            VSMCX_HIP(hipMemcpyAsync(replication, f->rb.rep_out, 8 * N, hipMemcpyDeviceToHost, s));
        std::vector<uint32_t> anc(N);
        VSMCX_HIP(hipMemcpyAsync(anc.data(), f->rb.anc, 4 * N, hipMemcpyDeviceToHost, s));
        FilterCtl h2;
        VSMCX_HIP(hipMemcpyAsync(&h2, f->ctl, sizeof(h2), hipMemcpyDeviceToHost, s));
        std::vector<double> us, tm;
        const bool sorted = f->d.scheme == kMultinomial || f->d.scheme == kResidual;
        if (uniforms && sorted) {
            us.resize(N);
            tm.resize((N + kFoldTile - 1) / kFoldTile + 1);
            VSMCX_HIP(hipMemcpyAsync(us.data(), f->rb.uscan, 8 * N, hipMemcpyDeviceToHost, s));
            VSMCX_HIP(hipMemcpyAsync(tm.data(), f->rb.u_tmaxex, 8 * tm.size(),
                                     hipMemcpyDeviceToHost, s));
        }
        c->sync();
        check_dev_err(h2);
        if (copy_from)
            for (uint64_t i = 0; i < N; ++i) copy_from[i] = anc[i];
        if (nn) *nn = h2.nn;
        if (uniforms && sorted)
            for (uint64_t k = 0; k < h2.nn; ++k)
                uniforms[k] = us[k] < tm[k / kFoldTile] ? tm[k / kFoldTile] : us[k];
    } else {
        c->sync();
    }
    c->release(f->rb.rep_out);
    f->rb.rep_out = nullptr;
    f->cur ^= 1;
    ++f->t;
    API_END
}

}  // extern "C"
