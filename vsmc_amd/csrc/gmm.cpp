// gmm.cpp -- the C ABI of the Gaussian-mixture SMC sampler with adaptive
// tempering (vsmcx_gmm_*): the reference's gmm_smc driver
// (example/gmm/src/gmm_smc.cpp:43-52 with example/include/smc_do.hpp:57-66:
// initialize, then iterate while alpha < 1) over device-resident kernels.
//
// Host code only.  One iteration is one stream of launches with no host
// synchronisation inside it; after each iteration the host reads back the
// iteration's record (one row, pinned) to see whether alpha has reached 1.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"

using namespace vsmcx;
using namespace vsmcx_host;

static_assert(sizeof(GmmRow) == sizeof(vsmcx_gmm_step), "history row layout");

struct vsmcx_gmm {
    vsmcx_ctx *ctx = nullptr;
    vsmcx_gmm_desc d{};
    bool fp64 = false;
    int K = 0;
    uint64_t N = 0, ndata = 0, tiles = 0, mh_blocks = 0;
    uint32_t passes = 0;
    GmmModel g{};
    double *xd = nullptr;
    float *xf = nullptr;
    void *par[2] = {nullptr, nullptr};
    double *llh[2] = {nullptr, nullptr};
    int cur = 0;
    double *lw = nullptr, *W = nullptr, *part = nullptr, *ppart = nullptr;
    uint32_t *apart = nullptr;
    FilterCtl *ctl = nullptr;
    GmmCtl *gc = nullptr;
    GmmRow *hist = nullptr;
    ResampleBuffers rb{};
    GmmRow *row_h = nullptr;  // pinned
    uint64_t rows = 0;        // history rows written (0 before initialize)
    double alpha = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double ms[3] = {0, 0, 0};  // device time: temper + reweight, resample, MH sweeps

    void release_all()
    {
        void *ps[] = {xd, xf, par[0], par[1], llh[0], llh[1], lw, W, part, ppart, apart, ctl, gc,
                      hist};
        for (void *p : ps) ctx->release(p);
        free_resample(ctx, rb);
        if (row_h) (void)hipHostFree(row_h);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    void read_row(uint64_t t)
    {
        VSMCX_HIP(hipMemcpyAsync(row_h, hist + t, sizeof(GmmRow), hipMemcpyDeviceToHost, ctx->stream));
        VSMCX_HIP(hipStreamSynchronize(ctx->stream));
        FilterCtl h{};
        VSMCX_HIP(hipMemcpy(&h, ctl, sizeof(h), hipMemcpyDeviceToHost));
        if (h.err & kErrDegenerate)
            throw DevError(VSMCX_ERR_DEGENERATE, "the weights sum to zero or are not finite");
        if (h.err)
            throw DevError(VSMCX_ERR_INTERNAL, "device consistency check failed (bits " +
                                                   std::to_string(h.err) + ")");
        alpha = row_h->alpha;
        rows = t + 1;
    }
};

int vsmcx_gmm_create(vsmcx_ctx *ctx, const vsmcx_gmm_desc *d, const double *data, uint64_t ndata,
                     vsmcx_gmm **out)
{
    API_BEGIN
    REQUIRE(ctx && d && data && out, "NULL argument");
    REQUIRE(d->precision == VSMCX_MIXED || d->precision == VSMCX_FP64_FUSED,
            "the sampler runs the mixed or the fused FP64 pipeline");
    REQUIRE(d->components >= 1 && d->components <= (uint32_t)kGmmMaxK, "components must be 1 .. 6");
    REQUIRE(d->n >= 1 && d->n <= (1ull << 32) - 1, "particle count out of range");
    REQUIRE(ndata >= 2 && ndata <= (1ull << 31), "need at least two data points");
    REQUIRE(d->sweeps >= 1 && d->sweeps <= 1000, "sweeps must be 1 .. 1000");
    REQUIRE(d->cess_drop > 0 && d->cess_drop < 1, "cess_drop must lie in (0, 1)");
    REQUIRE(d->scheme >= VSMCX_MULTINOMIAL && d->scheme <= VSMCX_RESIDUAL_SYSTEMATIC,
            "unknown resampling scheme");
    REQUIRE(d->max_steps >= 1, "max_steps must be positive");
    double xmin = data[0], xmax = data[0];
    for (uint64_t i = 0; i < ndata; ++i) {
        REQUIRE(std::isfinite(data[i]), "data must be finite");
        xmin = data[i] < xmin ? data[i] : xmin;
        xmax = data[i] > xmax ? data[i] : xmax;
    }
    REQUIRE(xmax > xmin, "data must not be constant");
    std::unique_ptr<vsmcx_gmm> S(new vsmcx_gmm);
    S->ctx = ctx;
    S->d = *d;
    S->fp64 = d->precision == VSMCX_FP64_FUSED;
    S->K = (int)d->components;
    S->N = d->n;
    S->ndata = ndata;
    S->tiles = (S->N + kMoveTile - 1) / kMoveTile;
    S->mh_blocks = (S->N + kBlock - 1) / kBlock;
    // the search halves [l, u] from width <= 1 after its first step down to eps
    const double eps = std::max(1e-10, 1e-3 * d->cess_drop);
    S->passes = 2 + (uint32_t)std::ceil(std::log2(1.05 / eps));
    // prior constants from the data range (example/gmm/include/gmm_state.hpp:145-160)
    const double kappa = 1 / ((xmax - xmin) * (xmax - xmin));
    GmmModel &g = S->g;
    g.mu0 = 0.5 * (xmin + xmax);
    g.sd0 = 1 / std::sqrt(kappa);
    g.shape0 = 2;
    g.scale0 = 50 * kappa;
    g.n = (uint32_t)ndata;
    g.sweeps = d->sweeps;
    g.seed = d->seed;
    g.N = S->N;
    const size_t rb = S->fp64 ? 8 : 4;
    try {
        S->xd = ctx->alloc_n<double>(ndata);
        S->xf = ctx->alloc_n<float>(ndata);
        std::vector<float> xf(data, data + ndata);
        VSMCX_HIP(hipMemcpy(S->xd, data, 8 * ndata, hipMemcpyHostToDevice));
        VSMCX_HIP(hipMemcpy(S->xf, xf.data(), 4 * ndata, hipMemcpyHostToDevice));
        g.xd = S->xd;
        g.xf = S->xf;
        for (int k = 0; k < 2; ++k) {
            S->par[k] = ctx->alloc(rb * 3 * S->K * S->N);
            S->llh[k] = ctx->alloc_n<double>(S->N);
        }
        S->lw = ctx->alloc_n<double>(S->N);
        S->W = ctx->alloc_n<double>(S->N);
        S->part = ctx->alloc_n<double>(3 * std::max(S->tiles, S->mh_blocks));
        S->ppart = ctx->alloc_n<double>(S->mh_blocks);
        S->apart = ctx->alloc_n<uint32_t>(3 * S->mh_blocks);
        S->ctl = ctx->alloc_n<FilterCtl>(1);
        S->gc = ctx->alloc_n<GmmCtl>(1);
        S->hist = ctx->alloc_zero<GmmRow>(d->max_steps + 1);
        alloc_resample(ctx, S->rb, S->N, S->N, true);
        VSMCX_HIP(hipHostMalloc((void **)&S->row_h, sizeof(GmmRow), hipHostMallocDefault));
        for (hipEvent_t &e : S->ev) VSMCX_HIP(hipEventCreate(&e));
        FilterCtl h{};
        h.n = S->N;
        h.threshold = d->threshold;
        h.scheme = d->scheme;
        h.eng.key[0] = d->resample_seed;
        h.eng.index = 4;
        GmmCtl c{};
        c.target = 1 - d->cess_drop;
        c.eps = eps;
        c.llc = -0.5 * (double)ndata * 1.837877066409345483560659472811235279723;
        VSMCX_HIP(hipMemcpy(S->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
        VSMCX_HIP(hipMemcpy(S->gc, &c, sizeof(c), hipMemcpyHostToDevice));
        ctx->sync();
    } catch (...) {
        S->release_all();
        throw;
    }
    *out = S.release();
    API_END
}

int vsmcx_gmm_destroy(vsmcx_gmm *s)
{
    API_BEGIN
    if (!s) return VSMCX_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    s->release_all();
    delete s;
    API_END
}

int vsmcx_gmm_initialize(vsmcx_gmm *s)
{
    API_BEGIN
    REQUIRE(s, "NULL sampler");
    REQUIRE(s->rows == 0, "the sampler is already initialised");
    hipStream_t st = s->ctx->stream;
    launch_gmm_init(s->fp64, s->K, s->g, s->ctl, s->gc, s->par[0], s->llh[0], s->part, s->hist, st);
    VSMCX_HIP(hipGetLastError());
    s->cur = 0;
    s->read_row(0);
    API_END
}

int vsmcx_gmm_iterate(vsmcx_gmm *s, uint64_t max_iter, uint64_t *done, double *alpha)
{
    API_BEGIN
    REQUIRE(s, "NULL sampler");
    REQUIRE(s->rows >= 1, "initialise the sampler first");
    hipStream_t st = s->ctx->stream;
    uint64_t k = 0;
    while (k < max_iter && s->alpha < 1) {
        REQUIRE(s->rows <= s->d.max_steps, "history full: max_steps iterations done");
        const int c = s->cur;
        VSMCX_HIP(hipEventRecord(s->ev[0], st));
        launch_gmm_temper(s->ctl, s->gc, s->llh[c], s->lw, s->W, s->N, s->passes, s->part, s->hist,
                          s->d.scheme, st);
        VSMCX_HIP(hipEventRecord(s->ev[1], st));
        launch_resample_from_log_weights(s->ctl, s->lw, s->N, s->d.scheme, s->rb, nullptr, st);
        VSMCX_HIP(hipEventRecord(s->ev[2], st));
        launch_gmm_mh(s->fp64, s->K, s->g, s->ctl, s->gc, s->par[c], s->par[1 - c], s->llh[c],
                      s->llh[1 - c], s->rb.anc, s->lw, s->ppart, s->apart, s->hist, st);
        VSMCX_HIP(hipEventRecord(s->ev[3], st));
        VSMCX_HIP(hipGetLastError());
        s->cur = 1 - c;
        s->read_row(s->rows);
        for (int j = 0; j < 3; ++j) {
            float f = 0;
            VSMCX_HIP(hipEventElapsedTime(&f, s->ev[j], s->ev[j + 1]));
            s->ms[j] += f;
        }
        ++k;
    }
    if (done) *done = k;
    if (alpha) *alpha = s->alpha;
    API_END
}

int vsmcx_gmm_times(vsmcx_gmm *s, double *temper_ms, double *resample_ms, double *mh_ms)
{
    API_BEGIN
    REQUIRE(s, "NULL sampler");
    if (temper_ms) *temper_ms = s->ms[0];
    if (resample_ms) *resample_ms = s->ms[1];
    if (mh_ms) *mh_ms = s->ms[2];
    API_END
}

int vsmcx_gmm_rows(vsmcx_gmm *s, uint64_t *rows)
{
    API_BEGIN
    REQUIRE(s && rows, "NULL argument");
    *rows = s->rows;
    API_END
}

int vsmcx_gmm_history(vsmcx_gmm *s, uint64_t count, vsmcx_gmm_step *out)
{
    API_BEGIN
    REQUIRE(s && out, "NULL argument");
    REQUIRE(count <= s->rows, "count beyond the rows written");
    VSMCX_HIP(hipMemcpyAsync(out, s->hist, sizeof(GmmRow) * count, hipMemcpyDeviceToHost,
                             s->ctx->stream));
    s->ctx->sync();
    API_END
}

int vsmcx_gmm_state(vsmcx_gmm *s, double *params, double *llh, double *lw)
{
    API_BEGIN
    REQUIRE(s, "NULL sampler");
    REQUIRE(s->rows >= 1, "initialise the sampler first");
    const uint64_t N = s->N, P = 3 * (uint64_t)s->K;
    const int c = s->cur;
    s->ctx->sync();
    if (params) {
        std::vector<double> soa(P * N);
        if (s->fp64) {
            VSMCX_HIP(hipMemcpy(soa.data(), s->par[c], 8 * P * N, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(P * N);
            VSMCX_HIP(hipMemcpy(f.data(), s->par[c], 4 * P * N, hipMemcpyDeviceToHost));
            for (uint64_t j = 0; j < P * N; ++j) soa[j] = f[j];
        }
        for (uint64_t i = 0; i < N; ++i)
            for (uint64_t p = 0; p < P; ++p) params[i * P + p] = soa[p * N + i];
    }
    if (llh) VSMCX_HIP(hipMemcpy(llh, s->llh[c], 8 * N, hipMemcpyDeviceToHost));
    if (lw) {
        // the weight set: log weights, 0 for every particle while the weights
        // are equal (after initialize or a resample)
        FilterCtl h{};
        VSMCX_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
        if (h.prev_resampled || s->rows == 1) {
            std::memset(lw, 0, 8 * N);
        } else {
            VSMCX_HIP(hipMemcpy(lw, s->lw, 8 * N, hipMemcpyDeviceToHost));
        }
    }
    API_END
}
