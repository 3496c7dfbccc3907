// noalloc.cpp -- vsmcx_resample_noalloc: vsmcx_resample with every device
// buffer carved from one fixed array, `scratch`, so the call itself allocates
// no device memory.
//
// The array is a file-scope __device__ variable: every GPU the library runs on
// holds its own copy, at its own address, from when the code object loads until
// the process ends.  In host code the array's name is only a stand-in, so the
// device address is looked up once per device with hipGetSymbolAddress (from
// vsmcx_ctx_create, the device current) and used wherever host code passes the
// memory on.  Calls on one device share that device's array, so calls that
// overlap on one GPU (from two contexts, or two host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vsmcx.h"
#include "host.hpp"
#include "internal.hpp"
#include "launch.hpp"
#include "resample_layout.hpp"

namespace {

constexpr uint64_t kMax = VSMCX_NOALLOC_MAX;

// Everything one call uses: the resampling pipeline's buffers for (M, N),
// the replication output, the control block, the staged weights and the
// widened copy_from.
template <class A>
constexpr void layout_noalloc(A &a, ResampleBuffers &rb, FilterCtl *&ctl, double *&w,
                              uint64_t *&cf, uint64_t M, uint64_t N, bool lists)
{
    layout_resample(a, rb, M, N, lists);
    rb.rep_out = a.template n<uint64_t>(M);
    ctl = a.template n<FilterCtl>(1);
    w = a.template n<double>(M);
    cf = lists ? a.template n<uint64_t>(M) : nullptr;
}

constexpr uint64_t noalloc_bytes(uint64_t M, uint64_t N)
{
    LayoutCounter c;
    ResampleBuffers rb;
    FilterCtl *ctl = nullptr;
    double *w = nullptr;
    uint64_t *cf = nullptr;
    layout_noalloc(c, rb, ctl, w, cf, M, N, true);
    return c.used;
}

constexpr uint64_t kScratchBytes = noalloc_bytes(kMax, kMax);
static_assert(kScratchBytes < (64ull << 20), "the fixed scratch array stays modest");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// bump allocator over one device's scratch array
struct Carve {
    unsigned char *base;
    uint64_t cap, used;
    template <class T>
    T *n(uint64_t count)
    {
        const uint64_t b = layout_round((count ? count : 1) * sizeof(T));
        if (used + b > cap) throw std::runtime_error("vsmcx_resample_noalloc: scratch layout overflow");
        T *p = reinterpret_cast<T *>(base + used);
        used += b;
        return p;
    }
};

std::mutex g_mu;
std::vector<void *> g_ptr;  // per device ordinal

}  // namespace

namespace vsmcx_host {

void *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((size_t)device >= g_ptr.size()) g_ptr.resize(device + 1, nullptr);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

}  // namespace vsmcx_host

extern "C" int vsmcx_resample_noalloc(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                                      vsmcx_threefry4x64_state *eng, const double *weight,
                                      uint64_t *replication, uint64_t *copy_from)
{
    API_BEGIN
    if (!resample_args(ctx, scheme, M, N, eng, weight, replication)) return VSMCX_OK;
    REQUIRE(M <= kMax && N <= kMax,
            "vsmcx_resample_noalloc: M and N must be at most VSMCX_NOALLOC_MAX (2^18) "
            "particles, the size of its fixed scratch array; use vsmcx_resample");
    REQUIRE(ctx->scratch, "vsmcx_resample_noalloc: no scratch array on this device");
    const bool lists = copy_from && M == N;
    Carve c{static_cast<unsigned char *>(ctx->scratch), kScratchBytes, 0};
    ResampleBuffers rb;
    FilterCtl *ctl = nullptr;
    double *w = nullptr;
    uint64_t *cf = nullptr;
    layout_noalloc(c, rb, ctl, w, cf, M, N, lists);
    // every buffer is written before it is read, and the control block is
    // overwritten from the host, so nothing carried over from an earlier call
    // is read
    resample_on(ctx, scheme, M, N, eng, weight, replication, copy_from, rb, ctl, w, cf);
    API_END
}
