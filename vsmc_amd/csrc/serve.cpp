// serve.cpp -- the serving loop (vsmcx_serve_*): filtering jobs that arrive
// over time, admitted against an HBM budget, kept resident on the device
// across steps, and advanced with continuous batching: every tick steps all
// jobs that have an observation waiting, each phase one grid launch over the
// tiles of all of them (kernels_tick.hip); jobs admitted while a tick runs
// join the next one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"

namespace {

double now_us()
{
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

// First-fit allocator over [0, cap) in whole units, coalescing on free.
class RangeAlloc {
  public:
    explicit RangeAlloc(uint64_t cap = 0) { reset(cap); }
    void reset(uint64_t cap)
    {
        free_.clear();
        if (cap) free_[0] = cap;
        cap_ = cap;
        used_ = 0;
    }
    bool alloc(uint64_t n, uint64_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, start = it->first + n;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n)
    {
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t cap_ = 0, used_ = 0;
};

struct Job {
    vsmcx_job_desc d{};
    int state = 0;  // 0 queued, 1 running, 2 done
    int slot = -1;
    uint64_t done = 0;      // steps completed
    uint64_t issued = 0;    // steps launched (done plus the ticks in flight)
    uint64_t received = 0;  // observations received
    uint64_t uploaded = 0;  // observations copied to the device
    uint64_t p_off = 0, p_units = 0, s_off = 0, s_units = 0, o_off = 0, o_units = 0;
    std::vector<double> obs;        // every observation received (host copy)
    std::vector<double> arrival;    // arrival time of each observation (us)
    std::vector<StepRow> history;   // filled when the job finishes
};

constexpr uint64_t kUnit = kFoldTile;  // particle-arena unit

uint64_t align16(uint64_t b) { return (b + 15) & ~(uint64_t)15; }

// Up to two ticks are in flight on the stream: while the device runs one, the
// host stages and launches the next.  Each has its own staging buffers.
constexpr int kTickDepth = 2;

// One launched tick: its jobs, its staging buffer, and the finished jobs
// whose history rows it copies back (offset into that buffer's history area,
// or kNoRows when that area was full and the rows are read at completion).
struct Tick {
    std::vector<uint32_t> slots;
    int buf = 0;
    std::vector<std::pair<uint32_t, uint64_t>> finished;  // (slot, first row)
};
constexpr uint64_t kNoRows = ~0ull;

}  // namespace

struct vsmcx_serve {
    vsmcx_ctx *ctx = nullptr;
    bool fp64 = false;
    int scheme = kSystematic;
    double threshold = 0.5;
    CvParams cv{};
    LgParams lg{};
    uint32_t max_jobs = 0;
    uint64_t max_horizon = 0;
    RangeAlloc parts, states, obsr;
    uint64_t cap_units = 0, cap_state = 0, cap_obs = 0;
    BatchPool P{};
    BatchFilter *df = nullptr;
    std::vector<TickItem> items_h;
    std::vector<int> free_slots;
    std::vector<uint64_t> slot_job;
    // one tick's upload, staged on the host and sent as one copy
    std::vector<SlotInit> st_init;
    std::vector<ObsRun> st_runs;
    std::vector<AvailSet> st_avail;
    std::vector<double> st_pool;
    std::vector<uint8_t> st_mark;  // per slot: already in st_avail
    uint64_t pool_cap = 0, stage_cap = 0;
    // per staging buffer (one per tick in flight)
    uint8_t *stage_h[kTickDepth] = {};   // pinned
    uint8_t *stage_d[kTickDepth] = {};
    int *err_h[kTickDepth] = {};         // pinned: the device error bits after that tick
    StepRow *hist_h[kTickDepth] = {};    // pinned: history rows of jobs finishing in it
    uint64_t hist_cap = 0;               // rows per buffer
    hipEvent_t tev0[kTickDepth] = {}, tev1[kTickDepth] = {};
    std::deque<Tick> flight;             // launched, not yet completed (oldest first)
    int next_buf = 0;
    std::unordered_map<uint64_t, Job> jobs;
    std::deque<uint64_t> waiting;
    uint64_t next_id = 1;
    std::vector<uint32_t> tick;
    std::vector<double> lat_us;
    uint64_t steps = 0, updates = 0, jobs_done = 0;
    double busy_ms = 0, t_first = -1, t_last = 0;
    uint64_t peak_units = 0;

    int obs_dim(int model) const { return model == VSMCX_MODEL_LG ? lg.dim : 2; }
    int state_dim(int model) const { return model == VSMCX_MODEL_LG ? lg.dim : 4; }
    size_t real_bytes() const { return fp64 ? 8 : 4; }

    void release_all()
    {
        void *ps[] = {df, P.st0, P.st1, P.lw, P.lwc, P.anc, P.zl, P.sl, P.rep32, P.rep8,
                      P.spec, P.gval, P.ts, P.tagg, P.tpre, P.tzs, P.tsum, P.tzsx, P.tpart,
                      P.wx, P.thr, P.thr_fb,
                      stage_d[0], stage_d[1], P.ctl, P.hist, (void *)P.obs,
                      P.next_t, (void *)P.avail,
                      P.cur, P.err_any};
        for (void *q : ps) ctx->release(q);
        for (int b = 0; b < kTickDepth; ++b) {
            if (tev0[b]) (void)hipEventDestroy(tev0[b]);
            if (tev1[b]) (void)hipEventDestroy(tev1[b]);
            if (stage_h[b]) (void)hipHostFree(stage_h[b]);
            if (err_h[b]) (void)hipHostFree(err_h[b]);
            if (hist_h[b]) (void)hipHostFree(hist_h[b]);
        }
    }

    void mark_avail(const Job &J)
    {
        if (st_mark[J.slot]) {
            for (AvailSet &a : st_avail)
                if (a.slot == (uint32_t)J.slot) a.avail = J.uploaded;
            return;
        }
        st_mark[J.slot] = 1;
        st_avail.push_back(AvailSet{(uint32_t)J.slot, 0, J.uploaded});
    }

    // admit waiting jobs in arrival order while slots and memory last
    void admit()
    {
        hipStream_t s = ctx->stream;
        while (!waiting.empty() && !free_slots.empty()) {
            Job &J = jobs[waiting.front()];
            const uint64_t units = (J.d.n + kUnit - 1) / kUnit;
            const uint64_t sunits = units * kUnit * state_dim(J.d.model);
            const uint64_t ounits = J.d.horizon * obs_dim(J.d.model);
            uint64_t po, so, oo;
            if (!parts.alloc(units, &po)) return;
            if (!states.alloc(sunits, &so)) {
                parts.release(po, units);
                return;
            }
            if (!obsr.alloc(ounits, &oo)) {
                parts.release(po, units);
                states.release(so, sunits);
                return;
            }
            const int slot = free_slots.back();
            free_slots.pop_back();
            J.slot = slot;
            J.state = 1;
            J.p_off = po;
            J.p_units = units;
            J.s_off = so;
            J.s_units = sunits;
            J.o_off = oo;
            J.o_units = ounits;
            slot_job[slot] = waiting.front();
            waiting.pop_front();
            peak_units = std::max(peak_units, parts.used());
            SlotInit r{};
            r.slot = (uint32_t)slot;
            BatchFilter &F = r.f;
            F.n = J.d.n;
            F.horizon = J.d.horizon;
            F.off = po * kUnit;
            F.soff = so;
            F.obs_off = oo;
            F.hist_off = (uint64_t)slot * max_horizon;
            F.seed = J.d.seed;
            F.model = J.d.model == VSMCX_MODEL_LG ? kModelLg : kModelCv;
            FilterCtl &h = r.c;
            h.n = J.d.n;
            h.threshold = threshold;
            h.scheme = scheme;
            h.eng.key[0] = J.d.resample_seed;
            h.eng.index = 4;
            h.lwc = P.lwc ? P.lwc + F.off : nullptr;
            st_init.push_back(r);
            J.uploaded = 0;
            mark_avail(J);
            upload(J);
        }
    }

    // stage observations received but not yet on the device (whole
    // observations, as many as the staged pool holds; the rest go next tick)
    void upload(Job &J)
    {
        if (J.state != 1 || J.uploaded == J.received) return;
        const uint64_t D = obs_dim(J.d.model);
        const uint64_t room = (pool_cap - st_pool.size()) / D;
        const uint64_t k = std::min<uint64_t>(room, J.received - J.uploaded);
        if (!k) return;
        st_runs.push_back(ObsRun{J.o_off + D * J.uploaded, (uint32_t)st_pool.size(), (uint32_t)(D * k)});
        st_pool.insert(st_pool.end(), J.obs.begin() + D * J.uploaded, J.obs.begin() + D * (J.uploaded + k));
        J.uploaded += k;
        mark_avail(J);
    }

    // the oldest tick in flight has finished: account for it, retire the jobs
    // it finished (their history rows are on the host already)
    void complete()
    {
        const Tick k = flight.front();
        flight.pop_front();
        const double t = now_us();
        float ms = 0;
        VSMCX_HIP(hipEventElapsedTime(&ms, tev0[k.buf], tev1[k.buf]));
        busy_ms += ms;
        const int e = *err_h[k.buf];
        if (e & kErrDegenerate) throw DevError(VSMCX_ERR_DEGENERATE, "a job's weights sum to zero or are not finite");
        if (e) throw DevError(VSMCX_ERR_INTERNAL, "device consistency check failed (bits " + std::to_string(e) + ")");
        for (uint32_t slot : k.slots) {
            Job &J = jobs[slot_job[slot]];
            lat_us.push_back(t - J.arrival[J.done]);
            J.done += 1;
            steps += 1;
            updates += J.d.n;
        }
        for (const auto &fr : k.finished) {
            const uint32_t slot = fr.first;
            Job &J = jobs[slot_job[slot]];
            J.history.resize(J.d.horizon);
            if (fr.second != kNoRows) {
                std::memcpy(J.history.data(), hist_h[k.buf] + fr.second,
                            sizeof(StepRow) * J.d.horizon);
            } else {  // the tick's history area was full: read the rows now
                VSMCX_HIP(hipMemcpyAsync(J.history.data(), P.hist + (uint64_t)slot * max_horizon,
                                         sizeof(StepRow) * J.d.horizon, hipMemcpyDeviceToHost,
                                         ctx->stream));
                ctx->sync();
            }
            parts.release(J.p_off, J.p_units);
            states.release(J.s_off, J.s_units);
            obsr.release(J.o_off, J.o_units);
            free_slots.push_back(slot);
            slot_job[slot] = 0;
            J.state = 2;
            J.slot = -1;
            jobs_done += 1;
        }
        t_last = t;
    }

    // complete every tick in flight that the device has finished; with
    // `wait`, at least the oldest one
    void retire(bool wait)
    {
        while (!flight.empty()) {
            const int b = flight.front().buf;
            if (wait) {
                VSMCX_HIP(hipEventSynchronize(tev1[b]));
                wait = false;
            } else {
                const hipError_t q = hipEventQuery(tev1[b]);
                if (q == hipErrorNotReady) return;
                VSMCX_HIP(q);
            }
            complete();
        }
    }

    // one tick: the staged upload, then one step of every running job with an
    // observation on the device not yet stepped (nor in a tick in flight).
    // Returns the jobs stepped.  Needs a free staging buffer.
    uint64_t launch()
    {
        tick.clear();
        for (size_t slot = 0; slot < slot_job.size(); ++slot) {
            if (!slot_job[slot]) continue;
            const Job &J = jobs[slot_job[slot]];
            if (J.state == 1 && J.uploaded > J.issued) tick.push_back((uint32_t)slot);
        }
        items_h.clear();  // the jobs' first tiles, then their other tiles (TickItems)
        for (uint32_t slot : tick) items_h.push_back(TickItem{slot, 0});
        for (uint32_t slot : tick) {
            const uint64_t T = (jobs[slot_job[slot]].d.n + kUnit - 1) / kUnit;
            for (uint64_t tt = 1; tt < T; ++tt) items_h.push_back(TickItem{slot, (uint32_t)tt});
        }
        if (tick.empty() && st_init.empty() && st_runs.empty() && st_avail.empty()) return 0;
        const int b = next_buf;
        uint8_t *sh = stage_h[b], *sd = stage_d[b];
        // pack: init | runs | avail | jobs | items | pool, each 16-byte aligned
        uint64_t o = 0;
        const uint64_t o_init = o; o = align16(o + sizeof(SlotInit) * st_init.size());
        const uint64_t o_runs = o; o = align16(o + sizeof(ObsRun) * st_runs.size());
        const uint64_t o_av = o;   o = align16(o + sizeof(AvailSet) * st_avail.size());
        const uint64_t o_jobs = o; o = align16(o + 4 * tick.size());
        const uint64_t o_items = o; o = align16(o + sizeof(TickItem) * items_h.size());
        const uint64_t o_pool = o; o = align16(o + 8 * st_pool.size());
        if (o > stage_cap) throw DevError(VSMCX_ERR_INTERNAL, "tick upload exceeds its staging buffer");
        std::memcpy(sh + o_init, st_init.data(), sizeof(SlotInit) * st_init.size());
        std::memcpy(sh + o_runs, st_runs.data(), sizeof(ObsRun) * st_runs.size());
        std::memcpy(sh + o_av, st_avail.data(), sizeof(AvailSet) * st_avail.size());
        std::memcpy(sh + o_jobs, tick.data(), 4 * tick.size());
        std::memcpy(sh + o_items, items_h.data(), sizeof(TickItem) * items_h.size());
        std::memcpy(sh + o_pool, st_pool.data(), 8 * st_pool.size());
        hipStream_t s = ctx->stream;
        VSMCX_HIP(hipEventRecord(tev0[b], s));
        VSMCX_HIP(hipMemcpyAsync(sd, sh, o, hipMemcpyHostToDevice, s));
        TickUpload u{};
        u.init = (const SlotInit *)(sd + o_init);
        u.runs = (const ObsRun *)(sd + o_runs);
        u.avail = (const AvailSet *)(sd + o_av);
        u.pool = (const double *)(sd + o_pool);
        u.n_init = (uint32_t)st_init.size();
        u.n_runs = (uint32_t)st_runs.size();
        u.n_avail = (uint32_t)st_avail.size();
        launch_tick_upload(u, df, P, s);
        const TickItem *items_d = (const TickItem *)(sd + o_items);
        const TickItems items{items_d, items_d + tick.size(), (uint32_t)tick.size()};
        launch_tick(fp64, scheme, items, (uint32_t)items_h.size(),
                    (const uint32_t *)(sd + o_jobs), (uint32_t)tick.size(), df, P, cv, lg, s);
        VSMCX_HIP(hipGetLastError());
        // the history rows of the jobs this tick finishes, copied back behind it
        Tick k;
        k.buf = b;
        uint64_t rows = 0;
        for (uint32_t slot : tick) {
            Job &J = jobs[slot_job[slot]];
            J.issued += 1;
            if (J.issued < J.d.horizon) continue;
            uint64_t at = kNoRows;
            if (rows + J.d.horizon <= hist_cap) {
                at = rows;
                VSMCX_HIP(hipMemcpyAsync(hist_h[b] + at, P.hist + (uint64_t)slot * max_horizon,
                                         sizeof(StepRow) * J.d.horizon, hipMemcpyDeviceToHost, s));
                rows += J.d.horizon;
            }
            k.finished.push_back({slot, at});
        }
        VSMCX_HIP(hipMemcpyAsync(err_h[b], P.err_any, sizeof(int), hipMemcpyDeviceToHost, s));
        VSMCX_HIP(hipEventRecord(tev1[b], s));
        for (const AvailSet &a : st_avail) st_mark[a.slot] = 0;
        st_init.clear();
        st_runs.clear();
        st_avail.clear();
        st_pool.clear();
        k.slots = tick;
        flight.push_back(std::move(k));
        next_buf = (b + 1) % kTickDepth;
        if (t_first < 0) t_first = now_us();
        return tick.size();
    }
};

int vsmcx_serve_create(vsmcx_ctx *ctx, const vsmcx_serve_desc *d, vsmcx_serve **out)
{
    API_BEGIN
    REQUIRE(ctx && d && out, "NULL argument");
    REQUIRE(d->precision == VSMCX_MIXED || d->precision == VSMCX_FP64_FUSED,
            "the serving loop runs the mixed or the fused FP64 pipeline");
    REQUIRE(d->scheme == VSMCX_SYSTEMATIC || d->scheme == VSMCX_STRATIFIED,
            "the serving loop resamples with the systematic or the stratified scheme");
    REQUIRE(d->max_jobs >= 1 && d->max_horizon >= 1, "max_jobs and max_horizon must be positive");
    REQUIRE(d->lg_dim >= 1 && d->lg_dim <= (uint32_t)kLgMaxDim, "lg_dim must be 1 .. 10");
    REQUIRE(d->lg_a, "NULL lg_a");
    std::unique_ptr<vsmcx_serve> S(new vsmcx_serve);
    S->ctx = ctx;
    S->fp64 = d->precision == VSMCX_FP64_FUSED;
    S->scheme = d->scheme;
    S->threshold = d->threshold;
    S->cv = CvParams{d->dt, d->sd_pos, d->sd_vel, d->prior_sd_pos, d->prior_sd_vel, d->obs_scale,
                     d->nu, d->llh_const};
    LgParams &lg = S->lg;
    lg.dim = (int32_t)d->lg_dim;
    lg.q = d->lg_q;
    lg.r = d->lg_r;
    lg.s0 = d->lg_s0;
    lg.llh_const = -(double)d->lg_dim * std::log(d->lg_r * std::sqrt(2 * 3.14159265358979323846));
    for (uint32_t i = 0; i < d->lg_dim * d->lg_dim; ++i) {
        lg.ad[i] = d->lg_a[i];
        lg.a[i] = (float)d->lg_a[i];
    }
    S->max_jobs = d->max_jobs;
    S->max_horizon = d->max_horizon;
    // The budget covers the particle arena: per particle 80 B of weights,
    // ancestors, lists and fold scratch (84 B with FP32 state: the compact
    // log-weights) plus two states of the average model dimension
    // ((4 + lg_dim) / 2, the CV/LG mix).
    const double rb = S->real_bytes();
    const double avg_dim = 0.5 * (4 + d->lg_dim);
    const double per_particle = (rb == 4 ? 84 : 80) + 2 * rb * avg_dim;
    const uint64_t particles = (uint64_t)((double)d->hbm_budget / per_particle);
    S->cap_units = particles / kUnit;
    REQUIRE(S->cap_units >= 1, "HBM budget too small for one job");
    const uint64_t P_ = S->cap_units * kUnit;
    S->cap_state = (uint64_t)(P_ * avg_dim);
    S->cap_obs = (uint64_t)d->max_jobs * d->max_horizon * std::max<uint32_t>(2, d->lg_dim);
    S->parts.reset(S->cap_units);
    S->states.reset(S->cap_state);
    S->obsr.reset(S->cap_obs);
    try {
        BatchPool &P = S->P;
        P.st0 = ctx->alloc(S->cap_state * S->real_bytes());
        P.st1 = ctx->alloc(S->cap_state * S->real_bytes());
        P.lw = ctx->alloc_n<double>(P_);
        if (S->real_bytes() == 4) P.lwc = ctx->alloc_n<float>(P_);
        P.anc = ctx->alloc_n<uint32_t>(P_);
        P.zl = ctx->alloc_n<uint32_t>(P_);
        P.sl = ctx->alloc_n<uint32_t>(P_);
        P.rep32 = ctx->alloc_n<uint32_t>(P_);
        P.rep8 = ctx->alloc_n<uint8_t>(P_);
        P.spec = ctx->alloc_n<SpecRec>(P_);
        P.gval = ctx->alloc_n<double>(P_);
        P.ts = ctx->alloc_n<double>(S->cap_units);
        P.tagg = ctx->alloc_n<Seg>(S->cap_units);
        P.tpre = ctx->alloc_n<Seg>(S->cap_units);
        P.tzs = ctx->alloc_n<uint64_t>(S->cap_units);
        P.tsum = ctx->alloc_n<double>(S->cap_units);
        P.tzsx = ctx->alloc_n<uint64_t>(S->cap_units);
        P.tpart = ctx->alloc_n<double>(5 * S->cap_units);
        P.wx = ctx->alloc_n<double>(P_);
        P.thr = ctx->alloc_n<Seg>(S->cap_units * kBlock);
        P.thr_fb = ctx->alloc_n<int32_t>(S->cap_units * kBlock);
        // staging: every slot initialised, one run and one count per slot,
        // the job and item lists of a tick, and the observation pool
        S->pool_cap = std::max<uint64_t>(1 << 16, (uint64_t)d->max_jobs * std::max<uint32_t>(2, d->lg_dim) * 2);
        S->stage_cap = align16(sizeof(SlotInit) * d->max_jobs) + align16(sizeof(ObsRun) * d->max_jobs) +
                       align16(sizeof(AvailSet) * d->max_jobs) + align16(4 * (uint64_t)d->max_jobs) +
                       align16(sizeof(TickItem) * S->cap_units) + align16(8 * S->pool_cap);
        // history rows copied back per tick: 64 finishing jobs at the longest horizon
        S->hist_cap = 64 * d->max_horizon;
        for (int b = 0; b < kTickDepth; ++b) {
            S->stage_d[b] = (uint8_t *)ctx->alloc(S->stage_cap);
            VSMCX_HIP(hipHostMalloc((void **)&S->stage_h[b], S->stage_cap, hipHostMallocDefault));
            VSMCX_HIP(hipHostMalloc((void **)&S->err_h[b], sizeof(int), hipHostMallocDefault));
            *S->err_h[b] = 0;
            VSMCX_HIP(hipHostMalloc((void **)&S->hist_h[b], sizeof(StepRow) * S->hist_cap,
                                    hipHostMallocDefault));
            VSMCX_HIP(hipEventCreate(&S->tev0[b]));
            VSMCX_HIP(hipEventCreate(&S->tev1[b]));
        }
        P.ctl = ctx->alloc_n<FilterCtl>(d->max_jobs);
        P.hist = ctx->alloc_n<StepRow>((uint64_t)d->max_jobs * d->max_horizon);
        P.obs = ctx->alloc_n<double>(S->cap_obs);
        P.next_t = ctx->alloc_zero<uint64_t>(d->max_jobs);
        P.avail = ctx->alloc_zero<uint64_t>(d->max_jobs);
        P.cur = ctx->alloc_zero<uint32_t>(d->max_jobs);
        P.err_any = ctx->alloc_zero<int>(1);
        S->df = ctx->alloc_n<BatchFilter>(d->max_jobs);
        ctx->sync();
    } catch (...) {
        S->release_all();
        throw;
    }
    S->slot_job.assign(d->max_jobs, 0);
    S->st_mark.assign(d->max_jobs, 0);
    for (int i = (int)d->max_jobs - 1; i >= 0; --i) S->free_slots.push_back(i);
    *out = S.release();
    API_END
}

int vsmcx_serve_destroy(vsmcx_serve *s)
{
    API_BEGIN
    if (!s) return VSMCX_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    s->release_all();
    delete s;
    API_END
}

int vsmcx_serve_submit(vsmcx_serve *s, const vsmcx_job_desc *d, uint64_t *job)
{
    API_BEGIN
    REQUIRE(s && d && job, "NULL argument");
    REQUIRE(d->model == VSMCX_MODEL_CV || d->model == VSMCX_MODEL_LG, "unknown model");
    REQUIRE(d->n >= 1 && d->n <= (uint64_t)kBatchMaxTiles * kFoldTile,
            "jobs hold 1 to 2^20 particles");
    REQUIRE(d->horizon >= 1 && d->horizon <= s->max_horizon, "horizon out of range");
    const uint64_t id = s->next_id++;
    Job &J = s->jobs[id];
    J.d = *d;
    J.obs.reserve((size_t)d->horizon * s->obs_dim(d->model));
    s->waiting.push_back(id);
    *job = id;
    API_END
}

int vsmcx_serve_observe(vsmcx_serve *s, uint64_t job, const double *y)
{
    API_BEGIN
    REQUIRE(s && y, "NULL argument");
    auto it = s->jobs.find(job);
    REQUIRE(it != s->jobs.end(), "unknown job");
    Job &J = it->second;
    REQUIRE(J.received < J.d.horizon, "more observations than the job's horizon");
    const int D = s->obs_dim(J.d.model);
    J.obs.insert(J.obs.end(), y, y + D);
    J.arrival.push_back(now_us());
    J.received += 1;
    API_END
}

int vsmcx_serve_observe_many(vsmcx_serve *s, uint64_t count, const uint64_t *ids,
                             const double *const *ys)
{
    API_BEGIN
    REQUIRE(s && (count == 0 || (ids && ys)), "NULL argument");
    for (uint64_t k = 0; k < count; ++k) {
        const int rc = vsmcx_serve_observe(s, ids[k], ys[k]);
        if (rc != VSMCX_OK) return rc;
    }
    API_END
}

int vsmcx_serve_ready(vsmcx_serve *s, uint64_t max, uint64_t *ids, uint64_t *count)
{
    API_BEGIN
    REQUIRE(s && count && (max == 0 || ids), "NULL argument");
    uint64_t c = 0;
    for (const auto &kv : s->jobs) {
        const Job &J = kv.second;
        if (J.state == 2 || J.received > J.done || J.received >= J.d.horizon) continue;
        if (c < max) ids[c] = kv.first;
        ++c;
    }
    *count = c;
    API_END
}

int vsmcx_serve_poll(vsmcx_serve *s, int wait, uint64_t *launched)
{
    API_BEGIN
    REQUIRE(s, "NULL serve");
    if (launched) *launched = 0;
    // ticks the device has finished; with both staging buffers in use, wait
    // for the oldest (or return at once without wait)
    s->retire(false);
    if (s->flight.size() >= (size_t)kTickDepth) {
        if (!wait) return VSMCX_OK;
        s->retire(true);
    }
    s->admit();
    for (auto &kv : s->jobs)
        if (kv.second.state == 1) s->upload(kv.second);
    uint64_t n = s->launch();
    if (n == 0 && wait && !s->flight.empty()) {
        // nothing new to run: wait for the oldest tick, then try again (its
        // completion may free a slot or make a job ready)
        s->retire(true);
        s->admit();
        for (auto &kv : s->jobs)
            if (kv.second.state == 1) s->upload(kv.second);
        n = s->launch();
    }
    if (launched) *launched = n;
    API_END
}

int vsmcx_serve_job(vsmcx_serve *s, uint64_t job, int *state, uint64_t *steps)
{
    API_BEGIN
    REQUIRE(s, "NULL serve");
    auto it = s->jobs.find(job);
    REQUIRE(it != s->jobs.end(), "unknown job");
    if (state) *state = it->second.state;
    if (steps) *steps = it->second.done;
    API_END
}

int vsmcx_serve_history(vsmcx_serve *s, uint64_t job, uint64_t count, vsmcx_step_stats *out)
{
    API_BEGIN
    REQUIRE(s && out, "NULL argument");
    auto it = s->jobs.find(job);
    REQUIRE(it != s->jobs.end(), "unknown job");
    const Job &J = it->second;
    REQUIRE(J.state == 2, "the job has not finished");
    REQUIRE(count <= J.history.size(), "count beyond the job's horizon");
    std::memcpy(out, J.history.data(), sizeof(StepRow) * count);
    API_END
}

int vsmcx_serve_forget(vsmcx_serve *s, uint64_t job)
{
    API_BEGIN
    REQUIRE(s, "NULL serve");
    auto it = s->jobs.find(job);
    REQUIRE(it != s->jobs.end(), "unknown job");
    REQUIRE(it->second.state == 2, "only finished jobs can be forgotten");
    s->jobs.erase(it);
    API_END
}

int vsmcx_serve_stats(vsmcx_serve *s, vsmcx_serve_summary *out)
{
    API_BEGIN
    REQUIRE(s && out, "NULL argument");
    vsmcx_serve_summary r{};
    r.steps = s->steps;
    r.particle_updates = s->updates;
    r.jobs_done = s->jobs_done;
    uint64_t active = 0;
    for (const auto &kv : s->jobs) active += kv.second.state == 1;
    r.jobs_active = active;
    r.jobs_waiting = s->waiting.size();
    if (!s->lat_us.empty()) {
        std::vector<double> l = s->lat_us;
        const size_t i50 = (size_t)(0.50 * (l.size() - 1)), i99 = (size_t)(0.99 * (l.size() - 1));
        std::nth_element(l.begin(), l.begin() + i99, l.end());
        r.p99_us = l[i99];
        std::nth_element(l.begin(), l.begin() + i50, l.begin() + i99);
        r.p50_us = i50 == i99 ? r.p99_us : l[i50];
    }
    r.busy_ms = s->busy_ms;
    r.elapsed_ms = s->t_first < 0 ? 0 : (s->t_last - s->t_first) * 1e-3;
    r.peak_particles = s->peak_units * kUnit;
    r.capacity_particles = s->cap_units * kUnit;
    *out = r;
    API_END
}
