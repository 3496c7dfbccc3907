// batch.cpp -- the C ABI of batched filters (vsmcx_batch_*): pooled device
// memory for thousands of pf filters, one persistent kernel launch per run,
// and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <tuple>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"

struct vsmcx_batch {
    vsmcx_ctx *ctx = nullptr;
    uint32_t nf = 0;
    bool fp64 = false;
    int scheme = kSystematic;
    CvParams p{};
    LgParams lg{};
    std::vector<BatchFilter> hf;
    std::vector<uint64_t> avail_h;
    uint64_t total = 0, total_obs = 0;
    BatchFilter *df = nullptr;
    uint32_t *order_work = nullptr, *order_given = nullptr;
    BatchPool P{};
    unsigned *queue = nullptr;
    unsigned workgroups = 0;
    int order = 0;  // 0: most work first, 1: as given
    int strategy = VSMCX_BATCH_TICKS;
    std::vector<uint64_t> next_h;  // host mirror of each filter's next step
    uint32_t *jobs_dev = nullptr;  // ticks: filters by steps remaining, descending
    TickItem *items_dev = nullptr; // ticks: their tiles, in that order
    std::vector<uint32_t> jobs_h;
    std::vector<TickItem> items_h;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the tuner's candidates and their times (ms per tuning run)
    struct Cand {
        int strategy;
        unsigned wg;
        int order;
        double ms;
    };
    std::vector<Cand> tuned;

    size_t state_bytes() const { return fp64 ? 32 : 16; }

    // Everything a step reads and then overwrites, so that every tuning
    // candidate starts from the same state: both state buffers, log-weights,
    // ancestors, control blocks, step counters, buffer selectors, error bits
    // and the history rows.  The rest of the pool is scratch a step writes
    // before it reads.
    std::vector<std::pair<void *, size_t>> mutable_arrays() const
    {
        std::vector<std::pair<void *, size_t>> a = {{P.st0, total * state_bytes()},
                                                    {P.st1, total * state_bytes()},
                                                    {P.lw, 8 * total},
                                                    {P.anc, 4 * total},
                                                    {P.ctl, sizeof(FilterCtl) * nf},
                                                    {P.next_t, 8 * (size_t)nf},
                                                    {P.cur, 4 * (size_t)nf},
                                                    {P.err_any, sizeof(int)},
                                                    {P.hist, sizeof(StepRow) * total_obs}};
        if (P.lwc) a.push_back({P.lwc, 4 * total});
        return a;
    }
    void copy_arrays(const std::vector<void *> &saved, bool restore)
    {
        const auto arrs = mutable_arrays();
        for (size_t k = 0; k < arrs.size(); ++k) {
            void *live = arrs[k].first;
            VSMCX_HIP(hipMemcpyAsync(restore ? live : saved[k], restore ? saved[k] : live,
                                     arrs[k].second, hipMemcpyDeviceToDevice, ctx->stream));
        }
    }

    void release_all()
    {
        void *ps[] = {df, order_work, order_given, P.st0, P.st1, P.lw, P.lwc, P.anc, P.zl, P.sl,
                      P.rep32, P.rep8, P.spec, P.gval, P.ts, P.tagg, P.tpre, P.tzs, P.tsum, P.tzsx,
                      P.tpart, P.wx, P.thr, P.thr_fb, P.ctl, P.hist, (void *)P.obs, P.next_t,
                      (void *)P.avail, P.cur, P.err_any, queue, jobs_dev, items_dev};
        for (void *q : ps) ctx->release(q);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }

    // default persistent grid: the workgroups that fit at once on the device
    unsigned default_workgroups() const
    {
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
        const unsigned per_cu = scheme == kStratified ? 2 : 3;
        return std::min<unsigned>(nf, (unsigned)cus * per_cu);
    }

    // steps filter f takes in a run of at most max_steps
    uint64_t steps_of(uint32_t f, uint64_t max_steps) const
    {
        uint64_t t1 = std::min(avail_h[f], hf[f].horizon);
        if (t1 <= next_h[f]) return 0;
        return std::min(t1 - next_h[f], max_steps);
    }

    void launch(uint64_t max_steps, unsigned wg, int ord, int strat)
    {
        if (strat == VSMCX_BATCH_PERSISTENT) {
            launch_batch_run(fp64, scheme, df, ord == 0 ? order_work : order_given, nf, P, p, lg,
                             max_steps, queue, wg, ctx->stream);
            VSMCX_HIP(hipGetLastError());
        } else {
            launch_ticks(max_steps);
        }
        for (uint32_t f = 0; f < nf; ++f) next_h[f] += steps_of(f, max_steps);
    }

    // One tick per step: every filter with a step left in this run, each phase
    // one grid launch over all their tiles.  Sorted by steps left, the active
    // filters of tick k are a prefix of one list, uploaded once.
    void launch_ticks(uint64_t max_steps)
    {
        std::vector<std::pair<uint64_t, uint32_t>> act;
        for (uint32_t f = 0; f < nf; ++f) {
            const uint64_t sf = steps_of(f, max_steps);
            if (sf) act.push_back({sf, f});
        }
        if (act.empty()) return;
        std::stable_sort(act.begin(), act.end(),
                         [](const std::pair<uint64_t, uint32_t> &a,
                            const std::pair<uint64_t, uint32_t> &c) { return a.first > c.first; });
        jobs_h.clear();
        items_h.clear();
        // the jobs' first tiles, then their other tiles: the first j jobs'
        // items are the first j of the one and a prefix of the other
        std::vector<uint64_t> item_end;  // items of the first j jobs
        for (const auto &a : act) {
            jobs_h.push_back(a.second);
            items_h.push_back(TickItem{a.second, 0});
        }
        const size_t nfirst = items_h.size();
        uint64_t nall = 0;
        for (const auto &a : act) {
            const uint64_t T = (hf[a.second].n + kFoldTile - 1) / kFoldTile;
            for (uint64_t tt = 1; tt < T; ++tt) items_h.push_back(TickItem{a.second, (uint32_t)tt});
            nall += T;
            item_end.push_back(nall);
        }
        hipStream_t s = ctx->stream;
        VSMCX_HIP(hipMemcpyAsync(jobs_dev, jobs_h.data(), 4 * jobs_h.size(), hipMemcpyHostToDevice, s));
        VSMCX_HIP(hipMemcpyAsync(items_dev, items_h.data(), sizeof(TickItem) * items_h.size(),
                                 hipMemcpyHostToDevice, s));
        size_t nj = act.size();
        for (uint64_t k = 0; k < act.front().first; ++k) {
            while (nj > 0 && act[nj - 1].first <= k) --nj;  // filters done before tick k
            const TickItems items{items_dev, items_dev + nfirst, (uint32_t)nj};
            launch_tick(fp64, scheme, items, (uint32_t)item_end[nj - 1], jobs_dev, (uint32_t)nj,
                        df, P, p, lg, s, true);
        }
        VSMCX_HIP(hipGetLastError());
    }

    void check_errors()
    {
        int e = 0;
        VSMCX_HIP(hipMemcpyAsync(&e, P.err_any, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (e & kErrDegenerate) throw DevError(VSMCX_ERR_DEGENERATE, "a filter's weights sum to zero or are not finite");
        if (e) throw DevError(VSMCX_ERR_INTERNAL, "device consistency check failed (bits " + std::to_string(e) + ")");
    }
};

int vsmcx_batch_create(vsmcx_ctx *ctx, const vsmcx_batch_desc *d, vsmcx_batch **out)
{
    API_BEGIN
    REQUIRE(ctx && d && out, "NULL argument");
    REQUIRE(d->nfilters >= 1 && d->nfilters < (1ull << 31), "nfilters out of range");
    REQUIRE(d->n && d->horizon && d->seed && d->resample_seed, "NULL per-filter array");
    REQUIRE(d->precision == VSMCX_MIXED || d->precision == VSMCX_FP64_FUSED,
            "batched filters run the mixed or the fused FP64 pipeline");
    REQUIRE(d->scheme == VSMCX_SYSTEMATIC || d->scheme == VSMCX_STRATIFIED,
            "batched filters resample with the systematic or the stratified scheme");
    std::unique_ptr<vsmcx_batch> b(new vsmcx_batch);
    b->ctx = ctx;
    b->nf = (uint32_t)d->nfilters;
    b->fp64 = d->precision == VSMCX_FP64_FUSED;
    b->scheme = d->scheme;
    b->p = CvParams{d->dt, d->sd_pos, d->sd_vel, d->prior_sd_pos, d->prior_sd_vel, d->obs_scale,
                    d->nu, d->llh_const};
    b->hf.resize(b->nf);
    uint64_t off = 0, obs = 0;
    for (uint32_t f = 0; f < b->nf; ++f) {
        const uint64_t n = d->n[f], h = d->horizon[f];
        REQUIRE(n >= 1 && n <= (uint64_t)kBatchMaxTiles * kFoldTile,
                "batched filters hold 1 to 2^20 particles");
        REQUIRE(h >= 1, "horizon must be at least 1");
        BatchFilter &F = b->hf[f];
        F.n = n;
        F.horizon = h;
        F.off = off;
        F.soff = 4 * off;
        F.obs_off = 2 * obs;
        F.hist_off = obs;
        F.seed = d->seed[f];
        F.model = kModelCv;
        F.pad = 0;
        off += (n + kFoldTile - 1) / kFoldTile * kFoldTile;
        obs += h;
    }
    b->total = off;
    b->total_obs = obs;
    b->avail_h.assign(b->nf, 0);
    b->next_h.assign(b->nf, 0);
    try {
        BatchPool &P = b->P;
        P.st0 = ctx->alloc(off * b->state_bytes());
        P.st1 = ctx->alloc(off * b->state_bytes());
        P.lw = ctx->alloc_n<double>(off);
        if (!b->fp64) P.lwc = ctx->alloc_n<float>(off);
        P.anc = ctx->alloc_n<uint32_t>(off);
        P.zl = ctx->alloc_n<uint32_t>(off);
        P.sl = ctx->alloc_n<uint32_t>(off);
        P.rep32 = ctx->alloc_n<uint32_t>(off);
        P.rep8 = ctx->alloc_n<uint8_t>(off);
        P.spec = ctx->alloc_n<SpecRec>(off);
        P.gval = ctx->alloc_n<double>(off);
        P.ts = ctx->alloc_n<double>(off / kFoldTile);
        P.tagg = ctx->alloc_n<Seg>(off / kFoldTile);
        P.tpre = ctx->alloc_n<Seg>(off / kFoldTile);
        P.tzs = ctx->alloc_n<uint64_t>(off / kFoldTile);
        P.tsum = ctx->alloc_n<double>(off / kFoldTile);
        P.tzsx = ctx->alloc_n<uint64_t>(off / kFoldTile);
        P.tpart = ctx->alloc_n<double>(5 * (off / kFoldTile));
        P.wx = ctx->alloc_n<double>(off);
        P.thr = ctx->alloc_n<Seg>(off / kFoldTile * kBlock);
        P.thr_fb = ctx->alloc_n<int32_t>(off / kFoldTile * kBlock);
        b->jobs_dev = ctx->alloc_n<uint32_t>(b->nf);
        b->items_dev = ctx->alloc_n<TickItem>(off / kFoldTile);
        P.ctl = ctx->alloc_n<FilterCtl>(b->nf);
        P.hist = ctx->alloc_zero<StepRow>(obs);
        P.obs = ctx->alloc_zero<double>(2 * obs);
        P.next_t = ctx->alloc_zero<uint64_t>(b->nf);
        P.avail = ctx->alloc_zero<uint64_t>(b->nf);
        P.cur = ctx->alloc_zero<uint32_t>(b->nf);
        P.err_any = ctx->alloc_zero<int>(1);
        b->queue = ctx->alloc_zero<unsigned>(1);
        b->df = ctx->alloc_n<BatchFilter>(b->nf);
        b->order_work = ctx->alloc_n<uint32_t>(b->nf);
        b->order_given = ctx->alloc_n<uint32_t>(b->nf);
        std::vector<FilterCtl> hc(b->nf);
        for (uint32_t f = 0; f < b->nf; ++f) {
            FilterCtl h{};
            h.n = d->n[f];
            h.threshold = d->threshold;
            h.scheme = d->scheme;
            h.eng.key[0] = d->resample_seed[f];
            h.eng.index = 4;
            h.lwc = P.lwc ? P.lwc + b->hf[f].off : nullptr;
            hc[f] = h;
        }
        std::vector<uint32_t> given(b->nf), work(b->nf);
        std::iota(given.begin(), given.end(), 0u);
        work = given;
        std::stable_sort(work.begin(), work.end(), [&](uint32_t a, uint32_t c) {
            return b->hf[a].n * b->hf[a].horizon > b->hf[c].n * b->hf[c].horizon;
        });
        hipStream_t s = ctx->stream;
        VSMCX_HIP(hipMemcpyAsync(P.ctl, hc.data(), sizeof(FilterCtl) * b->nf, hipMemcpyHostToDevice, s));
        VSMCX_HIP(hipMemcpyAsync(b->df, b->hf.data(), sizeof(BatchFilter) * b->nf, hipMemcpyHostToDevice, s));
        VSMCX_HIP(hipMemcpyAsync(b->order_work, work.data(), 4 * b->nf, hipMemcpyHostToDevice, s));
        VSMCX_HIP(hipMemcpyAsync(b->order_given, given.data(), 4 * b->nf, hipMemcpyHostToDevice, s));
        VSMCX_HIP(hipEventCreate(&b->ev0));
        VSMCX_HIP(hipEventCreate(&b->ev1));
        ctx->sync();
    } catch (...) {
        b->release_all();
        throw;
    }
    b->workgroups = b->default_workgroups();
    *out = b.release();
    API_END
}

int vsmcx_batch_destroy(vsmcx_batch *b)
{
    API_BEGIN
    if (!b) return VSMCX_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    b->release_all();
    delete b;
    API_END
}

int vsmcx_batch_set_obs(vsmcx_batch *b, uint64_t f, uint64_t first, uint64_t count,
                        const double *obs)
{
    API_BEGIN
    REQUIRE(b && (obs || count == 0), "NULL argument");
    REQUIRE(f < b->nf, "filter index out of range");
    const BatchFilter &F = b->hf[f];
    REQUIRE(first + count <= F.horizon, "observations beyond the filter's horizon");
    if (count) {
        VSMCX_HIP(hipMemcpyAsync((double *)b->P.obs + F.obs_off + 2 * first, obs,
                                 16 * count, hipMemcpyHostToDevice, b->ctx->stream));
        if (first + count > b->avail_h[f]) {
            b->avail_h[f] = first + count;
            VSMCX_HIP(hipMemcpyAsync((uint64_t *)b->P.avail + f, &b->avail_h[f], 8,
                                     hipMemcpyHostToDevice, b->ctx->stream));
        }
        b->ctx->sync();
    }
    API_END
}

int vsmcx_batch_set_launch(vsmcx_batch *b, uint32_t workgroups, int order)
{
    API_BEGIN
    REQUIRE(b, "NULL batch");
    REQUIRE(order == 0 || order == 1, "order is 0 (most work first) or 1 (as given)");
    b->workgroups = workgroups ? std::min<uint32_t>(workgroups, b->nf) : b->default_workgroups();
    b->order = order;
    API_END
}

int vsmcx_batch_run(vsmcx_batch *b, uint64_t max_steps, double *elapsed_ms)
{
    API_BEGIN
    REQUIRE(b, "NULL batch");
    hipStream_t s = b->ctx->stream;
    VSMCX_HIP(hipEventRecord(b->ev0, s));
    b->launch(max_steps, b->workgroups, b->order, b->strategy);
    VSMCX_HIP(hipEventRecord(b->ev1, s));
    b->check_errors();
    if (elapsed_ms) {
        float ms = 0;
        VSMCX_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *elapsed_ms = ms;
    }
    API_END
}

int vsmcx_batch_tune(vsmcx_batch *b, uint64_t steps, uint32_t *workgroups, int *order)
{
    API_BEGIN
    REQUIRE(b && steps >= 1, "NULL batch or zero steps");
    const unsigned full = b->default_workgroups();
    // candidates: ticks, and persistent grids of 1/4 .. 4/4 of the resident
    // workgroups in either filter order
    std::vector<std::tuple<int, unsigned, int>> cand;
    cand.push_back({VSMCX_BATCH_TICKS, b->workgroups, b->order});
    for (unsigned k = 1; k <= 4; ++k) {
        const unsigned wg = std::max(1u, std::min<unsigned>(b->nf, full * k / 4));
        for (int ord = 0; ord < 2; ++ord) cand.push_back({VSMCX_BATCH_PERSISTENT, wg, ord});
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    hipStream_t s = b->ctx->stream;
    // every candidate runs the same `steps` steps from the same state, which is
    // restored afterwards: tuning advances no filter
    const auto arrs = b->mutable_arrays();
    std::vector<void *> saved;
    const std::vector<uint64_t> next_saved = b->next_h;
    try {
        for (const auto &a : arrs) saved.push_back(b->ctx->alloc(a.second));
        b->copy_arrays(saved, false);
        double best = 1e300;
        std::tuple<int, unsigned, int> pick{b->strategy, b->workgroups, b->order};
        b->tuned.clear();
        for (const auto &c : cand) {
            VSMCX_HIP(hipEventRecord(b->ev0, s));
            b->launch(steps, std::get<1>(c), std::get<2>(c), std::get<0>(c));
            VSMCX_HIP(hipEventRecord(b->ev1, s));
            b->check_errors();
            float ms = 0;
            VSMCX_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            b->tuned.push_back({std::get<0>(c), std::get<1>(c), std::get<2>(c), (double)ms});
            if (ms < best) {
                best = ms;
                pick = c;
            }
            b->copy_arrays(saved, true);
            b->next_h = next_saved;
        }
        b->ctx->sync();
        b->strategy = std::get<0>(pick);
        b->workgroups = std::get<1>(pick);
        b->order = std::get<2>(pick);
    } catch (...) {
        for (void *q : saved) b->ctx->release(q);
        throw;
    }
    for (void *q : saved) b->ctx->release(q);
    if (workgroups) *workgroups = b->workgroups;
    if (order) *order = b->order;
    API_END
}

int vsmcx_batch_tune_report(vsmcx_batch *b, uint64_t max, uint64_t *count, int *strategy,
                            uint32_t *workgroups, int *order, double *ms)
{
    API_BEGIN
    REQUIRE(b && count, "NULL argument");
    *count = b->tuned.size();
    for (uint64_t k = 0; k < std::min<uint64_t>(max, b->tuned.size()); ++k) {
        const auto &c = b->tuned[k];
        if (strategy) strategy[k] = c.strategy;
        if (workgroups) workgroups[k] = c.wg;
        if (order) order[k] = c.order;
        if (ms) ms[k] = c.ms;
    }
    API_END
}

int vsmcx_batch_set_strategy(vsmcx_batch *b, int strategy)
{
    API_BEGIN
    REQUIRE(b, "NULL batch");
    REQUIRE(strategy == VSMCX_BATCH_PERSISTENT || strategy == VSMCX_BATCH_TICKS,
            "strategy is VSMCX_BATCH_PERSISTENT or VSMCX_BATCH_TICKS");
    b->strategy = strategy;
    API_END
}

int vsmcx_batch_get_strategy(vsmcx_batch *b, int *strategy)
{
    API_BEGIN
    REQUIRE(b && strategy, "NULL argument");
    *strategy = b->strategy;
    API_END
}

int vsmcx_batch_steps(vsmcx_batch *b, uint64_t f, uint64_t *t)
{
    API_BEGIN
    REQUIRE(b && t, "NULL argument");
    REQUIRE(f < b->nf, "filter index out of range");
    VSMCX_HIP(hipMemcpyAsync(t, b->P.next_t + f, 8, hipMemcpyDeviceToHost, b->ctx->stream));
    b->ctx->sync();
    API_END
}

int vsmcx_batch_history(vsmcx_batch *b, uint64_t f, uint64_t count, vsmcx_step_stats *out)
{
    API_BEGIN
    REQUIRE(b && out, "NULL argument");
    REQUIRE(f < b->nf, "filter index out of range");
    uint64_t t = 0;
    VSMCX_HIP(hipMemcpyAsync(&t, b->P.next_t + f, 8, hipMemcpyDeviceToHost, b->ctx->stream));
    b->ctx->sync();
    REQUIRE(count <= t, "count beyond completed steps");
    VSMCX_HIP(hipMemcpyAsync(out, b->P.hist + b->hf[f].hist_off, sizeof(StepRow) * count,
                             hipMemcpyDeviceToHost, b->ctx->stream));
    b->ctx->sync();
    API_END
}

int vsmcx_batch_state(vsmcx_batch *b, uint64_t f, double *state, double *log_weight)
{
    API_BEGIN
    REQUIRE(b, "NULL batch");
    REQUIRE(f < b->nf, "filter index out of range");
    const BatchFilter &F = b->hf[f];
    const uint64_t N = F.n;
    hipStream_t s = b->ctx->stream;
    FilterCtl h;
    uint32_t cur = 0;
    VSMCX_HIP(hipMemcpyAsync(&h, b->P.ctl + f, sizeof(h), hipMemcpyDeviceToHost, s));
    VSMCX_HIP(hipMemcpyAsync(&cur, b->P.cur + f, 4, hipMemcpyDeviceToHost, s));
    b->ctx->sync();
    if (state) {
        const size_t sb = b->state_bytes();
        std::vector<unsigned char> raw(N * sb);
        std::vector<uint32_t> anc(N), sl(N);
        const void *src = (const unsigned char *)(cur ? b->P.st1 : b->P.st0) + F.soff * (sb / 4);
        VSMCX_HIP(hipMemcpyAsync(raw.data(), src, N * sb, hipMemcpyDeviceToHost, s));
        VSMCX_HIP(hipMemcpyAsync(anc.data(), b->P.anc + F.off, 4 * N, hipMemcpyDeviceToHost, s));
        VSMCX_HIP(hipMemcpyAsync(sl.data(), b->P.sl + F.off, 4 * N, hipMemcpyDeviceToHost, s));
        b->ctx->sync();
        for (uint64_t i = 0; i < N; ++i) {
            uint64_t j = i;
            if (h.prev_resampled) {
                const uint32_t a = anc[i];
                const uint64_t q = (a & kAncZero) ? (uint64_t)(a & ~kAncZero) : N;
                j = (a & kAncZero) ? (q < N ? sl[q] : i) : a;
                j = j < N ? j : i;
            }
            for (int c = 0; c < 4; ++c) {
                if (b->fp64) {
                    double v;
                    std::memcpy(&v, raw.data() + j * 32 + 8 * c, 8);
                    state[4 * i + c] = v;
                } else {
                    float v;
                    std::memcpy(&v, raw.data() + j * 16 + 4 * c, 4);
                    state[4 * i + c] = v;
                }
            }
        }
    }
    if (log_weight) {
        launch_lw_expand(b->P.ctl + f, b->P.lw + F.off, N, b->P.wx + F.off, s);
        VSMCX_HIP(hipMemcpyAsync(log_weight, b->P.wx + F.off, 8 * N, hipMemcpyDeviceToHost, s));
        b->ctx->sync();
        for (uint64_t i = 0; i < N; ++i)
            log_weight[i] = h.prev_resampled ? 0.0 : log_weight[i] - h.m;
    }
    API_END
}

int vsmcx_batch_particles(vsmcx_batch *b, uint64_t *total)
{
    API_BEGIN
    REQUIRE(b && total, "NULL argument");
    uint64_t s = 0;
    for (const auto &F : b->hf) s += F.n;
    *total = s;
    API_END
}
