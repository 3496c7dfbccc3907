/*
 * vsmcx.h -- C ABI of the MI355X-native vSMC particle hot path.
 *
 * Plain C types only (pointers, sizes, fixed-width integers); every function
 * returns VSMCX_OK (0) or a negative error code and leaves a message for
 * vsmcx_last_error() (thread-local).  No exception crosses this boundary.
 * A context owns one HIP stream on one device and is not thread-safe: use
 * one context per host thread.
 *
 * Each entry names the vSMC interface it replaces (paths in the vSMC tree).
 */
#ifndef VSMCX_H
#define VSMCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define VSMCX_OK 0
#define VSMCX_ERR_ARG -1          /* invalid argument */
#define VSMCX_ERR_HIP -2          /* HIP runtime failure */
#define VSMCX_ERR_DEGENERATE -3   /* weights sum to zero or are not finite */
#define VSMCX_ERR_BUDGET -4       /* device-memory budget exceeded */
#define VSMCX_ERR_INTERNAL -5     /* device-side consistency check failed */
#define VSMCX_ERR_STATE -6        /* call out of sequence */

/* Resampling schemes: same values as vsmc::ResampleScheme
 * (include/vsmc/resample/common.hpp:124-131). */
#define VSMCX_MULTINOMIAL 0
#define VSMCX_RESIDUAL 1
#define VSMCX_STRATIFIED 2
#define VSMCX_SYSTEMATIC 3
#define VSMCX_RESIDUAL_STRATIFIED 4
#define VSMCX_RESIDUAL_SYSTEMATIC 5

/* Precision / pipeline of a filter. */
#define VSMCX_MIXED 0       /* FP32 state + likelihood, FP64 normaliser/scan/log Z, fused */
#define VSMCX_FP64_FUSED 1  /* FP64 everywhere, fused kernels */
#define VSMCX_FP64_PLAIN 2  /* FP64, one kernel per vSMC operation (baseline) */

const char *vsmcx_last_error(void);
const char *vsmcx_version(void);

/* ---------------------------------------------------------------- context */
typedef struct vsmcx_ctx vsmcx_ctx;
/* device ordinal; hbm_budget bytes this context may allocate (0 = no limit) */
int vsmcx_ctx_create(int device, uint64_t hbm_budget, vsmcx_ctx **out);
int vsmcx_ctx_destroy(vsmcx_ctx *ctx);
int vsmcx_ctx_sync(vsmcx_ctx *ctx);
/* bytes currently allocated by the context */
int vsmcx_ctx_allocated(vsmcx_ctx *ctx, uint64_t *bytes);
/* Streaming rates of the device, a measurement aid with no counterpart in the
 * reference: a copy of `bytes` bytes from one buffer to another, a read of
 * them and a write of them (16 B per lane, grid-stride), each the best of
 * `reps` timed launches; gbps[0..2] = copy (read + write bytes), read, write
 * in 10^9 bytes/s.  Allocates 2 x bytes from the context's budget for the
 * call.  bytes in [2^20, 2^36], reps in [1, 100]. */
int vsmcx_stream_probe(vsmcx_ctx *ctx, uint64_t bytes, int reps, double gbps[3]);

/* ---------------------------------------------------------------- RNG
 * Bulk block generation: block k (k = 0 .. nblocks-1) is the engine's
 * block at counter ctr0 + 1 + k, i.e. what nblocks*4 successive
 * operator() calls return after eng.ctr(ctr0) (include/vsmc/rng/philox.hpp:
 * 374-395, include/vsmc/rng/threefry.hpp:343-364).  `out` is host memory of
 * 4*nblocks words.  Keys are the engine key() words (Threefry: key(), not par). */
int vsmcx_philox4x32_fill(vsmcx_ctx *ctx, const uint32_t key[2], const uint32_t ctr0[4],
                          uint64_t nblocks, uint32_t *out);
int vsmcx_threefry4x64_fill(vsmcx_ctx *ctx, const uint64_t key[4], const uint64_t ctr0[4],
                            uint64_t nblocks, uint64_t *out);
int vsmcx_threefry4x32_fill(vsmcx_ctx *ctx, const uint32_t key[4], const uint32_t ctr0[4],
                            uint64_t nblocks, uint32_t *out);

/* Threefry4x64 engine state as the resample engine holds it
 * (include/vsmc/rng/threefry.hpp:512-516: ctr_, par_ (first 4 words = key),
 * buffer_, index_). */
typedef struct {
    uint64_t ctr[4];
    uint64_t key[4];
    uint64_t buffer[4];
    uint64_t index;
} vsmcx_threefry4x64_state;

/* ---------------------------------------------------------------- resample
 * vsmcx_resample replaces the resample operator vSMC stores as
 *   std::function<void(size_t M, size_t N, Threefry4x64 &rng,
 *                      const double *weight, size_t *replication)>
 * (include/vsmc/core/particle.hpp:61-62; Resample<Scheme>::operator() in
 * include/vsmc/resample/{multinomial,residual,...}.hpp:53-77).  Host pointers, synchronous.  `eng`
 * is advanced by exactly the draws vSMC makes.  `replication` receives M
 * counts summing to N.  When copy_from is not NULL (requires M == N) it also
 * receives internal::cfrp_trans's layout (include/vsmc/resample/common.hpp:
 * 87-118), i.e. what Particle::resample (core/particle.hpp:230-250) uses. */
int vsmcx_resample(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                   vsmcx_threefry4x64_state *eng, const double *weight,
                   uint64_t *replication, uint64_t *copy_from);

/* vsmcx_resample_noalloc: no counterpart in vSMC.  It is vsmcx_resample (same
 * arguments, same results bit for bit) for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): it calls no hipMalloc / hipFree.  Every device buffer it
 * uses is carved from one fixed array per GPU (about 26 MB, held from library
 * load until the process ends), so M and N must be at most VSMCX_NOALLOC_MAX
 * (VSMCX_ERR_ARG beyond it).  Calls on one GPU share that array: calls that
 * overlap on one GPU (two contexts or two host threads) race. */
#define VSMCX_NOALLOC_MAX (1ull << 18)
int vsmcx_resample_noalloc(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                           vsmcx_threefry4x64_state *eng, const double *weight,
                           uint64_t *replication, uint64_t *copy_from);

/* Exact left fold out[i] = fl(out[i-1] + x[i]) of non-negative x, computed on
 * the device (the running sum inside include/vsmc/resample/common.hpp:76-77
 * and math::asum, include/vsmc/math/cblas.hpp:53-63).  Host pointers. */
int vsmcx_fold_prefix(vsmcx_ctx *ctx, uint64_t n, const double *x, double *out);

/* ---------------------------------------------------------------- filter
 * Device-resident bootstrap particle filter for the constant-velocity model
 * of example/pf (state x, y, vx, vy; Student-t position likelihood).  One
 * step replays Sampler::do_iter (include/vsmc/core/sampler.hpp:783-834):
 * move + log-likelihood -> WeightSet::add_log_weight (core/weight_set.hpp:
 * 324-331) -> ESS test and Particle::resample (core/particle.hpp:230-250)
 * -> monitor at MonitorMCMC stage.  Step 0 is the init step
 * (core/sampler.hpp:773-781). */
typedef struct vsmcx_filter vsmcx_filter;

typedef struct {
    uint64_t n;            /* particles */
    uint64_t horizon;      /* observation capacity (steps) */
    int precision;         /* VSMCX_MIXED / VSMCX_FP64_FUSED / VSMCX_FP64_PLAIN */
    int scheme;            /* VSMCX_* resampling scheme */
    double threshold;      /* resample when ESS < threshold * n */
    uint64_t seed;         /* Philox4x32 key of the move stream */
    uint64_t resample_seed;/* Threefry4x64 seed of the resample engine */
    /* model */
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel;
    double obs_scale, nu, llh_const; /* llh_const: per-coordinate log normaliser */
} vsmcx_filter_desc;

typedef struct {
    double ess;       /* ESS before the resampling decision */
    double log_z;     /* accumulated log-evidence */
    double mean_x;    /* filter mean at MonitorMCMC stage */
    double mean_y;
    int32_t resampled;
    int32_t pad;
    uint64_t draws;   /* resample-engine draws this step */
} vsmcx_step_stats;

int vsmcx_filter_create(vsmcx_ctx *ctx, const vsmcx_filter_desc *desc, vsmcx_filter **out);
int vsmcx_filter_destroy(vsmcx_filter *f);
/* observations: T x 2 doubles (x, y), T <= horizon */
int vsmcx_filter_set_obs(vsmcx_filter *f, uint64_t T, const double *obs);
/* run steps [t0, t0 + nsteps): enqueued on the context stream, no host sync */
int vsmcx_filter_run(vsmcx_filter *f, uint64_t nsteps);
/* finish: evaluate the monitor of the last step, synchronise, check errors */
int vsmcx_filter_finish(vsmcx_filter *f);
/* Checkpoint / resume (on-disk state; the reference's hdf5store, utility/hdf5io.hpp:
 * 610-746, writes a Sampler's state and summary to HDF5, which this image lacks): a
 * fused filter's particles, log-weights, ancestors, weight scalars, log-evidence,
 * resample-engine state, observations and history in one file of this build's own
 * versioned layout.  vsmcx_filter_load creates a filter on ctx that resumes exactly
 * where the save left it: later steps equal an uninterrupted run's bit for bit. */
int vsmcx_filter_save(vsmcx_filter *f, const char *path);
/* the descriptor the filter was created with (or loaded from) */
int vsmcx_filter_describe(vsmcx_filter *f, vsmcx_filter_desc *out);
int vsmcx_filter_load(vsmcx_ctx *ctx, const char *path, vsmcx_filter **out);
/* steps completed so far */
int vsmcx_filter_steps(vsmcx_filter *f, uint64_t *t);
/* history rows [0, count) (after finish) */
int vsmcx_filter_history(vsmcx_filter *f, uint64_t count, vsmcx_step_stats *out);
/* current state as n x 4 doubles (x, y, vx, vy), log-weights (n), resample
 * engine; the state is the post-resample state of the last completed step */
int vsmcx_filter_state(vsmcx_filter *f, double *state, double *log_weight);
int vsmcx_filter_engine(vsmcx_filter *f, vsmcx_threefry4x64_state *eng);

/* Timing of the run's device work with HIP events on the context stream:
 * runs nsteps steps and reports elapsed milliseconds, and the average
 * duration of the fused move kernel (or the plain move kernel). */
int vsmcx_filter_time(vsmcx_filter *f, uint64_t nsteps, double *elapsed_ms,
                      double *move_kernel_ms);

/* ---------------------------------------------------------------- timing
 * HIP events on the context stream.  A timer holds n marks; elapsed()
 * synchronises on the later mark.  Profiling a filter brackets every move
 * kernel launch with an event pair; move_times() returns their durations in
 * launch order and forgets them.  No vSMC counterpart (vSMC times whole runs
 * with StopWatch, include/vsmc/utility/stop_watch.hpp:117-221). */
typedef struct vsmcx_timer vsmcx_timer;
int vsmcx_timer_create(vsmcx_ctx *ctx, uint64_t n, vsmcx_timer **out);
int vsmcx_timer_record(vsmcx_timer *tm, uint64_t i);
int vsmcx_timer_elapsed(vsmcx_timer *tm, uint64_t i, uint64_t j, double *ms);
int vsmcx_timer_destroy(vsmcx_timer *tm);
int vsmcx_filter_profile(vsmcx_filter *f, int enable);
int vsmcx_filter_move_times(vsmcx_filter *f, uint64_t max, uint64_t *count, double *ms);

/* ---------------------------------------------------------------- debug
 * Hooks for parity tests: run one step but stop after the resampling
 * decision and expose the normalised weights W (n doubles) the resampler
 * reads; then resample and expose replication / copy_from.  Sorted-uniform
 * schemes also expose the uniform array the device inverted (nn values). */
int vsmcx_filter_debug_step(vsmcx_filter *f, double *weight, uint64_t *replication,
                            uint64_t *copy_from, double *uniforms, uint64_t *nn,
                            vsmcx_threefry4x64_state *eng_before, int *resampled);

/* ---- batched filters ----------------------------------------------------------
 * Thousands of independent pf filters (each the vsmcx_filter model with its own
 * particle count, horizon, observations and seeds) advanced together by
 * vsmcx_batch_run, as one persistent launch or as one tick per step (see the
 * strategies below).  A filter's results do not depend on the other filters
 * in the batch, on the strategy or on the launch shape: they are the bits
 * vsmcx_filter computes for the same filter.  No single counterpart in vSMC: it replaces one Sampler per filter
 * (include/vsmc/core/sampler.hpp:450-483, 773-834), each stepped in turn. */
typedef struct vsmcx_batch vsmcx_batch;

typedef struct {
    uint64_t nfilters;
    const uint64_t *n;             /* [nfilters] particles, 1 .. 2^20 */
    const uint64_t *horizon;       /* [nfilters] steps */
    const uint64_t *seed;          /* [nfilters] Philox4x32 keys of the move streams */
    const uint64_t *resample_seed; /* [nfilters] Threefry4x64 seeds of the resample engines */
    int precision;                 /* VSMCX_MIXED or VSMCX_FP64_FUSED */
    int scheme;                    /* VSMCX_SYSTEMATIC or VSMCX_STRATIFIED */
    double threshold;              /* resample when ESS < threshold * n */
    /* model, shared by all filters (vsmcx_filter_desc's fields) */
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel;
    double obs_scale, nu, llh_const;
} vsmcx_batch_desc;

int vsmcx_batch_create(vsmcx_ctx *ctx, const vsmcx_batch_desc *desc, vsmcx_batch **out);
int vsmcx_batch_destroy(vsmcx_batch *b);
/* observation rows [first, first + count) of filter f, (x, y) pairs */
int vsmcx_batch_set_obs(vsmcx_batch *b, uint64_t f, uint64_t first, uint64_t count,
                        const double *obs);
/* One launch: every filter advances by up to max_steps steps, as far as its
 * observations allow.  elapsed_ms (may be NULL) receives the device time. */
int vsmcx_batch_run(vsmcx_batch *b, uint64_t max_steps, double *elapsed_ms);
/* Launch shape: workgroups of the persistent grid (0 = default) and the order
 * filters are taken in (0 = most work first, 1 = as given). */
int vsmcx_batch_set_launch(vsmcx_batch *b, uint32_t workgroups, int order);
/* Execution strategy.  PERSISTENT: one launch, workgroups take filters from a
 * queue and run whole steps of one filter each.  TICKS: one step of every
 * filter with a step left per tick, each phase one grid launch over all their
 * tiles (the serving loop's tick).  Both run the same tile bodies in the same
 * orders, so the results are the same bits either way. */
#define VSMCX_BATCH_PERSISTENT 0
#define VSMCX_BATCH_TICKS 1
int vsmcx_batch_set_strategy(vsmcx_batch *b, int strategy);
int vsmcx_batch_get_strategy(vsmcx_batch *b, int *strategy);
/* Autotuner, after the OpenCL manager's local-size sweep (profile_kernel,
 * include/vsmc/opencl/cl_manager.hpp:385-441): every candidate (ticks, and
 * persistent grids of several sizes and both filter orders) runs the same
 * `steps` steps from the same state, which is restored after each, so tuning
 * advances no filter; the fastest is kept (the results do not depend on the
 * choice).  vsmcx_batch_tune_report gives the last tuning's candidates (up to
 * max) and their times in ms; *count receives how many there were. */
int vsmcx_batch_tune(vsmcx_batch *b, uint64_t steps, uint32_t *workgroups, int *order);
int vsmcx_batch_tune_report(vsmcx_batch *b, uint64_t max, uint64_t *count, int *strategy,
                            uint32_t *workgroups, int *order, double *ms);
int vsmcx_batch_steps(vsmcx_batch *b, uint64_t f, uint64_t *t);
int vsmcx_batch_history(vsmcx_batch *b, uint64_t f, uint64_t count, vsmcx_step_stats *out);
int vsmcx_batch_state(vsmcx_batch *b, uint64_t f, double *state, double *log_weight);
int vsmcx_batch_particles(vsmcx_batch *b, uint64_t *total);

/* ---- serving loop --------------------------------------------------------------
 * Filtering jobs arrive over time (vsmcx_serve_submit), each with its own model,
 * particle count, horizon and seeds; their observations arrive one step at a
 * time (vsmcx_serve_observe).  Jobs are admitted in arrival order while a job
 * slot and room in the HBM arena are free, keep their particle state on the
 * device until their last step, and are advanced with continuous batching:
 * each vsmcx_serve_poll completes the ticks the device has finished, admits
 * jobs, and launches one tick: one step of every running job with an
 * observation waiting, each phase of the step one grid launch over the tiles
 * of all those jobs.  Up to two ticks are in flight, so the host stages the
 * next tick while the device runs one; a job's steps stay in order.  A job's
 * results do not depend on the other jobs.  No counterpart in vSMC (one
 * Sampler per job, stepped by its caller). */
#define VSMCX_MODEL_CV 0  /* the pf example's constant-velocity model */
#define VSMCX_MODEL_LG 1  /* linear-Gaussian: x_t = A x_{t-1} + q z, y_t = x_t + r v */

typedef struct vsmcx_serve vsmcx_serve;

typedef struct {
    uint64_t hbm_budget;   /* bytes of device memory for the particle arena */
    uint32_t max_jobs;     /* job slots on the device */
    uint32_t pad;
    uint64_t max_horizon;  /* longest job */
    int precision;         /* VSMCX_MIXED or VSMCX_FP64_FUSED */
    int scheme;            /* VSMCX_SYSTEMATIC or VSMCX_STRATIFIED */
    double threshold;
    /* constant-velocity model (vsmcx_filter_desc's fields) */
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel, obs_scale, nu, llh_const;
    /* linear-Gaussian model: dimension (<= 10), row-major A, noise sds, prior sd */
    uint32_t lg_dim;
    uint32_t pad2;
    const double *lg_a;
    double lg_q, lg_r, lg_s0;
} vsmcx_serve_desc;

typedef struct {
    int model;             /* VSMCX_MODEL_CV or VSMCX_MODEL_LG */
    int pad;
    uint64_t n;            /* particles, 1 .. 2^20 */
    uint64_t horizon;      /* steps (observations) */
    uint64_t seed;         /* Philox4x32 key of the move stream */
    uint64_t resample_seed;/* Threefry4x64 seed of the resample engine */
} vsmcx_job_desc;

typedef struct {
    uint64_t steps;              /* job steps completed */
    uint64_t particle_updates;   /* sum of n over completed steps */
    uint64_t jobs_done, jobs_active, jobs_waiting;
    double p50_us, p99_us;       /* step latency: observation arrival to step completion */
    double busy_ms;              /* device time of the ticks */
    double elapsed_ms;           /* first launch to last completion */
    uint64_t peak_particles;     /* most particle slots in use at once */
    uint64_t capacity_particles; /* particle slots the budget allows */
} vsmcx_serve_summary;

int vsmcx_serve_create(vsmcx_ctx *ctx, const vsmcx_serve_desc *desc, vsmcx_serve **out);
int vsmcx_serve_destroy(vsmcx_serve *s);
int vsmcx_serve_submit(vsmcx_serve *s, const vsmcx_job_desc *job, uint64_t *id);
/* one step's observation: 2 values (CV) or lg_dim values (LG) */
int vsmcx_serve_observe(vsmcx_serve *s, uint64_t id, const double *y);
/* one observation each for `count` jobs (ys[k] holds job ids[k]'s values) */
int vsmcx_serve_observe_many(vsmcx_serve *s, uint64_t count, const uint64_t *ids,
                             const double *const *ys);
/* jobs (waiting or running) that have used every observation they received
 * and expect more: up to `max` ids, *count = how many there are */
int vsmcx_serve_ready(vsmcx_serve *s, uint64_t max, uint64_t *ids, uint64_t *count);
/* complete the ticks the device has finished; with two in flight, wait for
 * the oldest (`wait`) or return at once; admit jobs, launch the next tick
 * (*launched = job steps in it).  With `wait`, when nothing could be
 * launched, wait for the oldest tick in flight and try once more. */
int vsmcx_serve_poll(vsmcx_serve *s, int wait, uint64_t *launched);
/* state: 0 waiting, 1 running, 2 done */
int vsmcx_serve_job(vsmcx_serve *s, uint64_t id, int *state, uint64_t *steps);
int vsmcx_serve_history(vsmcx_serve *s, uint64_t id, uint64_t count, vsmcx_step_stats *out);
int vsmcx_serve_forget(vsmcx_serve *s, uint64_t id);
int vsmcx_serve_stats(vsmcx_serve *s, vsmcx_serve_summary *out);

/* ---- device-resident particle values ----------------------------------------------
 * Storage for vSMC's value type T of Particle<T> (include/vsmc/core/particle.hpp:
 * 64-70, 230-250) kept in HBM: n rows of dim values (FP32 for VSMCX_MIXED, FP64
 * for VSMCX_FP64_FUSED).  vsmcx_state_copy is StateMatrix::copy
 * (include/vsmc/core/state_matrix.hpp:285-300), x[to] = x[copy_from[to]], done
 * as a device gather; the C++ adaptor vsmcx::StateHIP (include/vsmcx.hpp)
 * wraps it as a vSMC value type. */
typedef struct vsmcx_state vsmcx_state;

int vsmcx_state_create(vsmcx_ctx *ctx, uint64_t n, uint32_t dim, int precision,
                       vsmcx_state **out);
int vsmcx_state_destroy(vsmcx_state *s);
/* a second state with the same rows (Particle<T> copies its value) */
int vsmcx_state_clone(vsmcx_state *s, vsmcx_state **out);
int vsmcx_state_size(vsmcx_state *s, uint64_t *n, uint32_t *dim);
/* rows: n x dim, row-major, FP64 on the host */
int vsmcx_state_write(vsmcx_state *s, const double *rows);
int vsmcx_state_read(vsmcx_state *s, double *rows);
/* the current rows on the device (changes after vsmcx_state_copy), for user kernels */
int vsmcx_state_data(vsmcx_state *s, void **device_rows);
/* StateMatrix::copy: n must equal the state's size; copy_from is a host array */
int vsmcx_state_copy(vsmcx_state *s, uint64_t n, const uint64_t *copy_from);
/* The pf example's move_state (example/pf/include/pf_smp.hpp:150-238) for every
 * particle of a 4-column state: the prior draw at iter 0, the transition after,
 * with the model constants and Philox key of `model`; inc (host, n values)
 * receives each particle's log-likelihood increment of observation obs_xy. */
int vsmcx_state_cv_move(vsmcx_state *s, const vsmcx_filter_desc *model, uint64_t iter,
                        const double *obs_xy, double *inc);

/* ---- device weight set: moves written as user kernels ------------------------------
 * vSMC's WeightSet (include/vsmc/core/weight_set.hpp:61-470) with its log-weights in
 * HBM, updated from device arrays -- the increments a user's own HIP kernel writes, as
 * MoveCL kernels do (include/vsmc/opencl/backend_cl.hpp:689-708; such kernels draw
 * randoms with include/vsmcx_device.hpp, the analogue of rng/urng.h and
 * rng/normal01.h) -- and Particle::resample (include/vsmc/core/particle.hpp:230-250)
 * on the device: the ESS test at `threshold`, the scheme's resampling by the exact fold
 * (vSMC's inversion and cfrp_trans), and the gather of a vsmcx_state in place.  The
 * Threefry4x64 resample engine is seeded with resample_seed.  set / add / resample
 * enqueue on the context's stream (vsmcx_ctx_stream) and return without waiting; a
 * user kernel launched on that stream is ordered with them.  Every set or add is one
 * history row (at most max_steps): ESS before resampling, log-evidence (log sum_i W_i
 * exp(inc_i) accumulated), resampled, engine draws. */
typedef struct vsmcx_weights vsmcx_weights;

/* the context's HIP stream (a hipStream_t) */
int vsmcx_ctx_stream(vsmcx_ctx *ctx, void **hip_stream);
int vsmcx_weights_create(vsmcx_ctx *ctx, uint64_t n, int scheme, double threshold,
                         uint64_t resample_seed, uint64_t max_steps, vsmcx_weights **out);
int vsmcx_weights_destroy(vsmcx_weights *w);
/* a device buffer of n doubles owned by w, for a user kernel's increments */
int vsmcx_weights_buffer(vsmcx_weights *w, double **inc_dev);
/* set_log_weight (weight_set.hpp:296-307): lw = inc; a new run, history row 0 */
int vsmcx_weights_set_log(vsmcx_weights *w, const double *inc_dev);
/* add_log_weight (weight_set.hpp:324-331): lw += inc (lw = inc after a resample) */
int vsmcx_weights_add_log(vsmcx_weights *w, const double *inc_dev);
/* Particle::resample(op, threshold): if the last set / add left ESS < threshold * n,
 * resample and gather state's rows in place (state may be NULL); else nothing */
int vsmcx_weights_resample(vsmcx_weights *w, vsmcx_state *state);
/* the ancestors of the last resample on the device: n uint32 (cfrp_trans copy_from) */
int vsmcx_weights_ancestors(vsmcx_weights *w, const uint32_t **anc_dev);
int vsmcx_weights_steps(vsmcx_weights *w, uint64_t *steps);
/* rows 0 .. count-1; waits for the device and checks its error bits */
int vsmcx_weights_history(vsmcx_weights *w, uint64_t count, vsmcx_step_stats *out);
/* read_weight (weight_set.hpp:197-201): the n normalised weights */
int vsmcx_weights_read(vsmcx_weights *w, double *weight);

/* ---- Gaussian-mixture SMC with adaptive tempering ------------------------------
 * The reference's gmm_smc sampler (example/gmm/src/gmm_smc.cpp:43-52 driven by
 * example/include/smc_do.hpp:57-66 under the "cess" schedule): particles carry
 * K means, precisions and weights; each iteration picks the next temperature
 * alpha by bisection so that the conditional ESS of the reweighting is
 * 1 - cess_drop (example/include/move_alpha.hpp:242-279), reweights by
 * alpha_inc * log-likelihood (example/include/move_smc.hpp:57-77), resamples
 * when ESS < threshold * n, and runs `sweeps` sweeps of the random-walk MH
 * moves on the means, the precisions and the weights
 * (example/gmm/include/gmm_move.hpp:35-121).  Iterations continue until
 * alpha = 1.  Records carry the SMC and the path-sampling log evidence
 * (example/include/smc_do.hpp:75-79, include/vsmc/core/path.hpp:229-239). */
typedef struct vsmcx_gmm vsmcx_gmm;

typedef struct {
    uint64_t n;            /* particles */
    uint32_t components;   /* K, 1 .. 6 */
    uint32_t sweeps;       /* MH sweeps per iteration, each: mu, lambda, weight moves */
    int precision;         /* VSMCX_MIXED (FP32 parameters and likelihood terms,
                              FP64 weights, search and evidence) or VSMCX_FP64_FUSED */
    int scheme;            /* resampling scheme (the reference sampler: VSMCX_STRATIFIED) */
    double threshold;      /* resample when ESS < threshold * n (0.5) */
    double cess_drop;      /* next alpha keeps CESS = 1 - cess_drop (0.1) */
    uint64_t seed;         /* Philox4x32 key of the initial draws and the MH moves */
    uint64_t resample_seed;/* Threefry4x64 seed of the resample engine */
    uint64_t max_steps;    /* iteration cap (history rows = max_steps + 1) */
} vsmcx_gmm_desc;

typedef struct {
    double alpha, alpha_inc;    /* temperature after the iteration and its increment */
    double ess;                 /* ESS after the reweighting (before any resample) */
    double log_z;               /* SMC log evidence so far (with the likelihood constant) */
    double path_integrand;      /* sum_i W_i log-likelihood_i after the MH sweeps */
    double path_log_z;          /* path-sampling log evidence so far */
    int32_t resampled;
    uint32_t bisections;        /* steps of the tempering search */
    uint64_t accept[3];         /* accepted mu, lambda, weight moves (all sweeps) */
} vsmcx_gmm_step;

/* data: ndata observations (the prior's location and scale come from their range) */
int vsmcx_gmm_create(vsmcx_ctx *ctx, const vsmcx_gmm_desc *desc, const double *data,
                     uint64_t ndata, vsmcx_gmm **out);
int vsmcx_gmm_destroy(vsmcx_gmm *g);
/* Sampler::initialize (include/vsmc/core/sampler.hpp:450-462): prior draws,
 * alpha = 0, equal weights; writes history row 0 */
int vsmcx_gmm_initialize(vsmcx_gmm *g);
/* up to max_iter iterations, stopping once alpha = 1; *done = iterations run,
 * *alpha = temperature after them (either may be NULL) */
int vsmcx_gmm_iterate(vsmcx_gmm *g, uint64_t max_iter, uint64_t *done, double *alpha);
int vsmcx_gmm_rows(vsmcx_gmm *g, uint64_t *rows);
/* device time of the iterations so far: tempering search with the reweighting,
 * resampling, and the MH sweeps (HIP events around each phase) */
int vsmcx_gmm_times(vsmcx_gmm *g, double *temper_ms, double *resample_ms, double *mh_ms);
int vsmcx_gmm_history(vsmcx_gmm *g, uint64_t count, vsmcx_gmm_step *out);
/* params: n x 3K row-major (K means, K precisions, K weights); llh: n; lw: n
 * unnormalised log weights (zero while the weights are equal); any may be NULL */
int vsmcx_gmm_state(vsmcx_gmm *g, double *params, double *llh, double *lw);

#ifdef __cplusplus
}
#endif
#endif /* VSMCX_H */
