// internal.hpp -- device-side control structures and launcher declarations
// shared by the HIP translation units and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "fold_core.hpp"
#include "rng_core.hpp"

namespace vsmcx {

constexpr int kBlock = 256;                  // threads per workgroup
constexpr int kItems = 16;                   // particles per thread
constexpr int kTile = kBlock * kItems;       // particles per workgroup tile
constexpr int kMoveItems = 8;                // particles per thread in the move kernels
constexpr int kMoveTile = kBlock * kMoveItems;  // particles per move workgroup (one partial)
constexpr int kFoldItems = 8;                // elements per thread in the fold kernels
constexpr int kFoldTile = kBlock * kFoldItems;  // elements per fold tile
constexpr int kFoldChunk = 64;                  // tiles per chunk of the segment scan
// the move's tile partials and the fold's per-tile arrays share one tiling
static_assert(kMoveTile == kFoldTile, "move and fold tiles coincide");

// Zero-slot marker of encoded ancestors: a slot that the resample left empty
// holds kAncZero | its zero rank, and its source is the surplus copy of that
// rank (cfrp_trans's pairing, include/vsmc/resample/common.hpp:91-110).  N <
// 2^31, so ranks fit below the marker.
constexpr uint32_t kAncZero = 0x80000000u;

// The source slot of slot i after a resample: anc[i], or, for an encoded zero
// slot, the surplus copy of its rank in sl.  Plain (decoded) ancestors pass
// through unchanged.
__host__ __device__ inline uint32_t anc_source(const uint32_t *anc, const uint32_t *sl, uint64_t i)
{
    const uint32_t a = anc[i];
    return (a & kAncZero) ? sl[a & ~kAncZero] : a;
}

// Resampling schemes, same order as vsmc::ResampleScheme
// (include/vsmc/resample/common.hpp:124-131).
enum Scheme : int {
    kMultinomial = 0,
    kResidual = 1,
    kStratified = 2,
    kSystematic = 3,
    kResidualStratified = 4,
    kResidualSystematic = 5
};

// error bits raised on the device
enum DevErr : int {
    kErrFoldClass = 1,       // fold classification bound violated
    kErrDegenerate = 2,      // weights sum to zero / NaN
    kErrWindow = 4,          // inversion window inconsistency
    kErrTimeout = 8,         // a look-back wait exceeded its spin limit
    kErrBounds = 16,         // a computed rank fell outside its buffer
    kErrSearch = 32,         // the tempering search did not end within its passes
};

// Per-filter control block, resident in device memory.  Kernels of one step
// communicate through it so the host never synchronises inside a step.
struct FilterCtl {
    double m;          // max log-weight
    double coeff;      // 1 / sum exp(lw - m)
    double lse;        // log sum exp(lw)
    double logz;       // accumulated log-evidence
    double ess;        // effective sample size of the current weights
    double rcoeff;     // residual schemes: 1 / fold(r)
    double u0;         // systematic: first uniform
    double delta;      // 1.0 / nn
    double fold_total; // last exact fold total
    double threshold;  // resample when ess < threshold * n
    uint64_t n;        // particles
    uint64_t nn;       // uniforms this resample (n, or n - R for residual)
    uint64_t r_int;    // residual integral total R
    uint64_t n_final;  // uniforms placed before the last bin
    uint64_t zeros;    // zero-replication slots this resample
    uint64_t t;        // iteration number of the step in flight
    uint64_t draws;    // engine draws of the last resample
    Tf64State eng;     // resample engine (Threefry4x64)
    int resample;      // this step resamples
    int prev_resampled;// the previous step resampled (move reads ancestors)
    int scheme;
    int err;
    uint64_t specials; // special elements of the last fold
    // Form of the current log-weights: 0 FP64 values in the lw buffer; 1
    // compact, the FP32 pipeline's log2 terms l2 (4 bytes each) in lwc with
    // lw = lw_c2 - lw_h l2, which the move stores on steps whose weights are the
    // increments alone (the first step and every step after a resample) when
    // the filter has an lwc buffer.  Two buffers, so a step that adds to
    // compact log-weights never writes over values it has yet to read.  The
    // move of the step in flight records its form in lw_store_kind; the
    // reduction after it makes that lw_kind.
    int lw_kind;
    int lw_store_kind;
    double lw_c2, lw_h;
    float *lwc;  // the filter's compact buffer (device), or null
};

// Log-weight i of the lw buffer in either form (FilterCtl::lw_kind); the
// compact form gives the very bits the move computed.
struct LwView {
    const double *lw;
    const float *lwc;
    int kind;
    double c2, h;
    __device__ __forceinline__ double operator()(uint64_t i) const
    {
        return kind ? c2 - h * (double)lwc[i] : lw[i];
    }
};
__device__ __forceinline__ LwView lw_view(const FilterCtl *ctl, const double *lw)
{
    return LwView{lw, ctl->lwc, ctl->lw_kind, ctl->lw_c2, ctl->lw_h};
}

// Parameters of the constant-velocity model (example/pf).
struct CvParams {
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel;
    double obs_scale, nu, llh_const;
};

// One row of per-step history.
struct StepRow {
    double ess, logz, mean_x, mean_y;
    int32_t resampled, pad;
    uint64_t draws;
};

// Scratch of the fold passes (kernels_fold.hip).
struct FoldScratch {
    double *tsum;      // [tiles] tile sums of x
    double *ts;        // [tiles] approximate exclusive prefix (fixed-order scan)
    Seg *agg;          // [tiles] segment summary of each tile
    Seg *pre;          // [tiles] exclusive segment summary within the tile's chunk of
                       // kFoldChunk tiles (compose with cpre of the chunk)
    Seg *cagg;         // [chunks] segment summary of each chunk
    Seg *cpre;         // [chunks] exclusive segment summary of the chunks
    uint32_t *tcnt;    // [tiles] pre[t].cnt: specials before tile t within its chunk
    SpecRec *spec;     // [tiles * kFoldTile] special elements, tile-local order
    double *gval;      // [M + 1] exact fold value after each special, fold order
    uint32_t *rep32;   // [M] residual integer parts; counts of 255 or more for the layout pass
    uint8_t *rep8;     // [M rounded up to whole tiles] replication counts for the layout
                       // pass, one byte each; 255: the count is in rep32
    uint64_t *zs;      // [tiles] zero slots << 32 | surplus copies
    uint64_t *zsx;     // [tiles] exclusive prefix of zs
    double *wx;        // [M] the values x_i, written by the tile-sum pass (optional)
    Seg *thr;          // [tiles * 256] each thread's exclusive segment prefix in its tile (optional)
    int32_t *thr_fb;   // [tiles * 256] per tile: the binade of a uniform tile (thr_store)
    size_t tiles;
};

struct HipError : std::runtime_error {
    explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define VSMCX_HIP(call)                                                               \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::vsmcx::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace vsmcx
