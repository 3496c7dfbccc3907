// resample_layout.hpp -- the device buffers of one resampling pipeline, laid out
// by any allocator with `template <class T> T *n(uint64_t count)`: the context's
// budgeted allocator (capi.cpp) or a bump allocator over a fixed array
// (noalloc.cpp).  The layout is constexpr-evaluable, so the fixed array's size
// is the byte count of this very layout at its maximum size.
#pragma once
#include <stdint.h>

#include "launch.hpp"

namespace vsmcx {

constexpr uint64_t kLayoutAlign = 256;

constexpr uint64_t layout_round(uint64_t b)
{
    return (b + kLayoutAlign - 1) / kLayoutAlign * kLayoutAlign;
}

// Fold / resample buffers for up to M particles and nuni sorted uniforms, with
// the ancestor and cfrp lists when `lists` is set, and with fold scratch of the
// uniforms' own when `own_ufs` is set (their fold then runs concurrently with
// the weight fold).
template <class A>
constexpr void layout_resample(A &a, ResampleBuffers &rb, uint64_t M, uint64_t nuni, bool lists,
                               bool own_ufs = false)
{
    // the fold scratch serves the weight fold (M elements) and the sorted
    // uniforms' fold (nuni elements)
    const uint64_t F = M > nuni ? M : nuni;
    const uint64_t tiles = (F + kFoldTile - 1) / kFoldTile;    // fold tiles
    const uint64_t utiles = (nuni + kFoldTile - 1) / kFoldTile; // fold tiles of the uniforms
    if (lists) {
        rb.anc = a.template n<uint32_t>(M);
        rb.zl = a.template n<uint32_t>(M);
        rb.sl = a.template n<uint32_t>(M);
    }
    rb.u_tmax = a.template n<double>(utiles + 1);
    rb.u_tmaxex = a.template n<double>(utiles + 1);
    rb.uscan = a.template n<double>(nuni ? nuni : 1);
    rb.nuni = nuni;
    FoldScratch &fs = rb.fs;
    fs.tsum = a.template n<double>(tiles);
    fs.ts = a.template n<double>(tiles);
    fs.agg = a.template n<Seg>(tiles);
    fs.pre = a.template n<Seg>(tiles);
    const uint64_t chunks = (tiles + kFoldChunk - 1) / kFoldChunk;
    fs.cagg = a.template n<Seg>(chunks);
    fs.cpre = a.template n<Seg>(chunks);
    fs.tcnt = a.template n<uint32_t>(tiles);
    fs.spec = a.template n<SpecRec>(tiles * kFoldTile);
    fs.gval = a.template n<double>(F + 1);
    fs.rep32 = a.template n<uint32_t>(M);
    fs.rep8 = a.template n<uint8_t>((M + kFoldTile - 1) / kFoldTile * kFoldTile);
    fs.zs = a.template n<uint64_t>(tiles);
    fs.zsx = a.template n<uint64_t>(tiles);
    fs.wx = a.template n<double>(M);
    fs.thr = a.template n<Seg>(tiles * kBlock);
    fs.thr_fb = a.template n<int32_t>(tiles * kBlock);
    fs.tiles = (M + kFoldTile - 1) / kFoldTile;
    FoldScratch &u = rb.ufs;
    u = fs;
    u.tsum = a.template n<double>(utiles + 1);
    u.ts = a.template n<double>(utiles + 1);
    u.wx = nullptr;
    u.tiles = utiles;
    if (own_ufs) {
        // every array the uniforms' fold writes (its inversion-free apply
        // writes no counts, lists or replication)
        const uint64_t uch = (utiles + kFoldChunk - 1) / kFoldChunk;
        u.agg = a.template n<Seg>(utiles);
        u.pre = a.template n<Seg>(utiles);
        u.cagg = a.template n<Seg>(uch);
        u.cpre = a.template n<Seg>(uch);
        u.tcnt = a.template n<uint32_t>(utiles);
        u.spec = a.template n<SpecRec>(utiles * kFoldTile);
        u.gval = a.template n<double>(nuni + 1);
        u.thr = a.template n<Seg>(utiles * kBlock);
        u.thr_fb = a.template n<int32_t>(utiles * kBlock);
        rb.uguess = a.template n<uint32_t>(tiles);
    }
}

// Counts the bytes a layout takes (every buffer rounded up to kLayoutAlign).
struct LayoutCounter {
    uint64_t used = 0;
    template <class T>
    constexpr T *n(uint64_t count)
    {
        used += layout_round((count ? count : 1) * sizeof(T));
        return nullptr;
    }
};

}  // namespace vsmcx
