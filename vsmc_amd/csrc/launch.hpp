// launch.hpp -- host-callable launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace vsmcx {

// Buffers of one resampling pipeline sized for M particles.
// The sorted uniforms' fold on a second stream of the filter, concurrent with
// the filter's own passes: it reads only the engine state and nn, which are
// fixed before the step's weights are (multinomial: from the start of the
// step; residual: once R is known).  It has its own control block (the fold
// passes write fold totals and error bits there) and its own fold scratch;
// the fork and join events order it against the filter's stream, and the
// weight fold's inversion waits for the join.  Owned by the filter.
struct SideFold {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t starts = nullptr;  // the weight fold's approximate tile starts written
    FilterCtl *uctl = nullptr;
    bool pending = false;  // the uniforms of the coming resample are already launched
};

struct ResampleBuffers {
    uint32_t *anc = nullptr;     // [M] ancestors (read by the next move)
    uint32_t *zl = nullptr;      // [M] zero-slot list
    uint32_t *sl = nullptr;      // [M] surplus-copy list
    uint64_t *rep_out = nullptr; // [M] optional replication output
    double *a_out = nullptr;     // [M] optional exact fold output
    double *uscan = nullptr;     // [nuni] sorted uniforms: running max within fold tiles
    double *u_tmax = nullptr;    // [nuni / kFoldTile] tile maxima
    double *u_tmaxex = nullptr;  // [nuni / kFoldTile] exclusive running max of tiles
    uint32_t *uguess = nullptr;  // [M / kFoldTile] guesses of the inversion's bounds (side fold)
    uint64_t nuni = 0;           // capacity of the uniform arrays
    FoldScratch fs{};
    // the sorted uniforms' own fold: tile sums and starts of its own; the
    // other arrays shared with fs when the two folds run one after the other,
    // its own when it runs on side.s
    FoldScratch ufs{};
    SideFold side{};
};


// kernels_fold.hip
// lse_part: the move's tile partials of lw ({max, sum exp(lw - max), ...},
// 5 doubles per 2048-particle tile), or nullptr when lw has none
// encode: the ancestors of zero slots are written encoded (kAncZero | zero
// rank, resolved through rb.sl by anc_source) and no zero-fill pass runs
void launch_resample_from_log_weights(FilterCtl *ctl, const double *lw, uint64_t M, int scheme,
                                      ResampleBuffers &rb, StepRow *hist, hipStream_t st,
                                      const double *lse_part = nullptr, bool pre_started = false,
                                      bool encode = false);
void launch_resample_from_weights(FilterCtl *ctl, const double *w, uint64_t M, int scheme,
                                  ResampleBuffers &rb, StepRow *hist, hipStream_t st);
void launch_fold_values(FilterCtl *ctl, const double *x, uint64_t M, ResampleBuffers &rb,
                        hipStream_t st);
// Multinomial with rb.side.s: the sorted uniforms of the step's resample on
// the side stream, forked from st now (before the move; the engine state is
// final once the previous step's work on st is done).  The resample then only
// waits for them.
void launch_side_uniforms(FilterCtl *ctl, ResampleBuffers &rb, hipStream_t st);
// out[i] = in[i] for i < n (uint32 -> uint64)
void launch_widen_u32(const uint32_t *in, uint64_t *out, uint64_t n, hipStream_t st);

// kernels_probe.hip: best ms of reps launches of a streaming probe (kind 0
// copy a -> b, 1 read a, 2 write b) over n float4 elements; negative on error
float launch_stream_probe(int kind, const float4 *a, float4 *b, uint64_t n, int reps,
                          hipStream_t st);

// kernels_rng.hip
void launch_philox4x32_fill(uint32_t k0, uint32_t k1, const uint32_t ctr0[4], uint64_t nblocks,
                            uint32_t *out, hipStream_t st);
void launch_threefry4x64_fill(const uint64_t key[4], const uint64_t ctr0[4], uint64_t nblocks,
                              uint64_t *out, hipStream_t st);
void launch_threefry4x32_fill(const uint32_t key[4], const uint32_t ctr0[4], uint64_t nblocks,
                              uint32_t *out, hipStream_t st);

// kernels_cv.hip
// fused move (+ init when t == 0), partials per tile: {max, sum e, sum e^2, mean x, mean y}
// anc / sl: the ancestors of the last resample and its surplus list (for
// encoded zero slots, see anc_source)
void launch_cv_move(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                    const uint32_t *anc, const uint32_t *sl, double *lw, const double *obs,
                    const CvParams &p, uint64_t seed, uint64_t t, uint64_t N, double *partials,
                    hipStream_t st);
// starts (optional, with prologue_scheme a non-residual scheme): on a resample
// the reduce also runs the scheme's prologue and writes the fold's tile starts
// of the filter's weights into starts->ts; the resample then runs with
// pre_started set
void launch_cv_reduce(FilterCtl *ctl, const double *partials, uint64_t ntiles, StepRow *hist,
                      uint64_t t, hipStream_t st, int prologue_scheme = -1,
                      const FoldScratch *starts = nullptr);
void launch_cv_monitor(bool fp64, FilterCtl *ctl, const void *state, const uint32_t *anc,
                       const uint32_t *sl, const double *lw, uint64_t N, double *partials,
                       StepRow *hist, hipStream_t st);
void launch_cv_weights(FilterCtl *ctl, const double *lw, uint64_t N, double *w, hipStream_t st);
// FP64 log-weights from the lw buffer in either form (FilterCtl::lw_kind)
void launch_lw_expand(const FilterCtl *ctl, const double *lw, uint64_t N, double *out,
                      hipStream_t st);
void launch_cv_gather(bool fp64, FilterCtl *ctl, const void *st_in, void *st_out,
                      const uint32_t *anc, const uint32_t *sl, uint64_t N, hipStream_t st);

// kernels_cv.hip: device-resident particle values (the vSMC value-type adaptor)
void launch_state_cv_move(bool fp64, void *st, double *inc, const CvParams &p, uint64_t seed,
                          uint64_t t, double ox, double oy, uint64_t N, hipStream_t s);
void launch_state_gather(bool fp64, const void *in, void *out, const uint64_t *copy_from,
                         uint32_t dim, uint64_t N, int *err, hipStream_t s);

// kernels_weights.hip (the device weight set)
void launch_ws_add(const FilterCtl *ctl, double *lw, const double *inc, uint64_t N, int set,
                   double *part, hipStream_t s);
void launch_ws_gather(bool fp64, const FilterCtl *ctl, void *x, const uint32_t *anc,
                      uint32_t dim, uint64_t N, int *err, hipStream_t s);
void launch_ws_done(FilterCtl *ctl, hipStream_t s);

// kernels_plain.hip: the un-fused FP64 baseline
void launch_plain_init_weights(double *w, uint64_t N, hipStream_t s);
void launch_plain_move(void *st, double *inc, const double *obs, const CvParams &p, uint64_t seed,
                       uint64_t t, uint64_t N, hipStream_t s);
void launch_plain_rest(FilterCtl *ctl, void *st, double *lw, double *w, double *inc,
                       double *part, double *scal, uint64_t t, uint64_t N, int scheme,
                       ResampleBuffers &rb, StepRow *hist, hipStream_t s);
uint64_t plain_partials(uint64_t N);


// ---- batched filters (kernels_batch.hip) --------------------------------------------
constexpr int kBatchMaxTiles = 512;  // filters of up to 512 * kFoldTile = 2^20 particles

// One filter of a batch (host-built, uploaded once).
struct BatchFilter {
    uint64_t n;         // particles
    uint64_t horizon;   // steps
    uint64_t off;       // first particle in the pooled per-particle arrays (a multiple of kFoldTile)
    uint64_t soff;      // first state element (Real units) in the pooled state arrays
    uint64_t obs_off;   // first observation element (doubles)
    uint64_t hist_off;  // first history row
    uint64_t seed;      // Philox4x32 key of the move stream
    int32_t model;      // kModelCv or kModelLg
    int32_t pad;
};

enum : int32_t { kModelCv = 0, kModelLg = 1 };

// Linear-Gaussian model: x_t = A x_{t-1} + q z, y_t = x_t + r v, x_0 ~ N(0, s0^2 I)
constexpr int kLgMaxDim = 10;
struct LgParams {
    int32_t dim;
    int32_t pad;
    double q, r, s0;
    double llh_const;   // -dim * log(r sqrt(2 pi))
    float a[kLgMaxDim * kLgMaxDim];   // row-major A (FP32 copy)
    double ad[kLgMaxDim * kLgMaxDim]; // row-major A (FP64)
};

// Pooled device arrays of a batch; per-particle arrays are indexed off + i,
// states soff + i * dim.
struct BatchPool {
    void *st0, *st1;       // states (float or double elements), ping-pong
    double *lw;
    float *lwc;            // compact log-weights (FP32 state only)
    uint32_t *anc, *zl, *sl, *rep32;
    uint8_t *rep8;         // replication counts as bytes (FoldScratch::rep8)
    SpecRec *spec;
    double *gval;
    double *ts;            // per tile (index off / kFoldTile + tile): approximate start
    Seg *tagg, *tpre;      // per tile: segment summary and exclusive prefix
    uint64_t *tzs;         // per tile: zeros << 32 | surplus (batched path: then their prefix)
    double *tsum;          // per tile (serving tick): sum of the weights
    uint64_t *tzsx;        // per tile (serving tick): exclusive prefix of tzs
    double *tpart;         // per tile (serving tick): 5 move partials
    double *wx;            // per particle (ticks): the weights the tile-sum pass wrote
    Seg *thr;              // per tile x 256 (ticks): each thread's segment prefix in its tile
    int32_t *thr_fb;       // per tile x 256 (ticks): the binade of a uniform tile (thr_store)
    FilterCtl *ctl;        // [filters]
    StepRow *hist;         // history rows
    const double *obs;     // observations (CV: 2 per step, LG: dim per step)
    uint64_t *next_t;      // [filters] next step
    const uint64_t *avail; // [filters] observations available
    uint32_t *cur;         // [filters] state buffer holding the current state
    int *err_any;          // OR of every filter's device error bits
};

// one (job slot, tile) work item of a serving tick
struct TickItem {
    uint32_t slot;
    uint32_t tile;
};

// A tick's work items in launch order: the first tile of every job, then the
// jobs' other tiles job by job.  A job's first tile holds most of its fold's
// binade crossings (the running sum starts near zero), so its workgroup takes
// the element-by-element path and runs longest; started first, these
// workgroups finish inside the launch instead of setting its tail.
struct TickItems {
    const TickItem *first;  // [nfirst] tile 0 of each job, in job order
    const TickItem *rest;   // the other tiles, job by job
    uint32_t nfirst;
    __device__ TickItem operator[](uint32_t b) const
    {
        return b < nfirst ? first[b] : rest[b - nfirst];
    }
};

// One tick's host-to-device traffic, packed into one pinned buffer and sent
// with one copy: slots (re)initialised for admitted jobs, runs of new
// observations, and the observation count now on the device per slot.
struct SlotInit {
    BatchFilter f;
    FilterCtl c;
    uint32_t slot;
    uint32_t pad[3];
};
struct ObsRun {
    uint64_t dst;    // first double in BatchPool::obs
    uint32_t src;    // first double in the staged pool
    uint32_t count;  // doubles
};
struct AvailSet {
    uint32_t slot;
    uint32_t pad;
    uint64_t avail;
};
struct TickUpload {
    const SlotInit *init;
    const ObsRun *runs;
    const AvailSet *avail;
    const double *pool;
    uint32_t n_init, n_runs, n_avail, pad;
};

// kernels_tick.hip: apply a staged upload (one launch; nothing if it is empty)
void launch_tick_upload(const TickUpload &u, BatchFilter *filters, const BatchPool &P,
                        hipStream_t st);

// kernels_tick.hip: one step of every listed job, one grid launch per phase
// cv_only: every listed job runs the CV model (the move then has no LG body)
void launch_tick(bool fp64, int scheme, const TickItems &items, uint32_t nitems,
                 const uint32_t *jobs, uint32_t njobs, const BatchFilter *filters,
                 const BatchPool &P, const CvParams &cv, const LgParams &lg, hipStream_t st,
                 bool cv_only = false);

void launch_batch_run(bool fp64, int scheme, const BatchFilter *filters, const uint32_t *order,
                      uint32_t nf, const BatchPool &P, const CvParams &p, const LgParams &lg,
                      uint64_t max_steps, unsigned *queue, unsigned grid, hipStream_t st);

// ---- Gaussian-mixture SMC with adaptive tempering (kernels_gmm.hip) --------------------
// Sampler state besides the weight set (FilterCtl), resident on the device.
struct GmmCtl {
    double alpha, alpha_inc;   // temperature and its last increment
    double l, u, a, a_old;     // tempering search (move_alpha.hpp:242-279)
    double target, eps;        // CESS target 1 - drop, search tolerance
    double zconst;             // sum of log-evidence increments
    double path, integrand;    // path-sampling sum and last integrand
    double llc;                // log-likelihood constant -n log(2 pi) / 2
    double sd[3];              // proposal scales of the mu, lambda, weight moves
    uint64_t t;                // iteration
    uint32_t done, iters;      // search finished; bisection steps taken
};

// One history row; the layout of vsmcx_gmm_step.
struct GmmRow {
    double alpha, alpha_inc, ess, log_z, path_integrand, path_log_z;
    int32_t resampled;
    uint32_t bisections;
    uint64_t accept[3];
};

constexpr int kGmmMaxK = 6;

// Constants of one sampler (host-built, passed by value).
struct GmmModel {
    double mu0, sd0, shape0, scale0;  // prior from the data range (gmm_state.hpp:145-160)
    const double *xd;                 // data, FP64
    const float *xf;                  // data, FP32 (mixed path)
    uint32_t n;                       // data points
    uint32_t sweeps;                  // MH sweeps per iteration
    uint64_t seed;                    // Philox key
    uint64_t N;                       // particles
};

void launch_gmm_init(bool fp64, int K, const GmmModel &g, FilterCtl *ctl, GmmCtl *gc, void *par,
                     double *llh, double *part, GmmRow *hist, hipStream_t st);
void launch_gmm_temper(FilterCtl *ctl, GmmCtl *gc, const double *llh, double *lw, double *W,
                       uint64_t N, uint32_t passes, double *part, GmmRow *hist, int scheme,
                       hipStream_t st);
void launch_gmm_mh(bool fp64, int K, const GmmModel &g, const FilterCtl *ctl, GmmCtl *gc,
                   const void *pin, void *pout, const double *llh_in, double *llh_out,
                   const uint32_t *anc, const double *lw, double *ppart, uint32_t *apart,
                   GmmRow *hist, hipStream_t st);

}  // namespace vsmcx
