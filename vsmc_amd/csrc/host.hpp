// host.hpp -- host-side plumbing shared by the C ABI translation units:
// error translation at the boundary and the context (device, stream, HBM
// budget, allocation bookkeeping).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>

#include "../../include/vsmcx.h"
#include "internal.hpp"

using namespace vsmcx;

namespace vsmcx_host {

inline thread_local std::string g_err;

struct BudgetError : std::runtime_error {
    explicit BudgetError(const std::string &s) : std::runtime_error(s) {}
};
struct ArgError : std::runtime_error {
    explicit ArgError(const std::string &s) : std::runtime_error(s) {}
};
struct DevError : std::runtime_error {
    int code;
    DevError(int c, const std::string &s) : std::runtime_error(s), code(c) {}
};

inline int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// the device's error bits of a control block, as the C ABI's status
inline void check_dev_err(const FilterCtl &h)
{
    if (h.err & kErrDegenerate) throw DevError(VSMCX_ERR_DEGENERATE, "weights sum to zero or are not finite");
    if (h.err) throw DevError(VSMCX_ERR_INTERNAL, "device consistency check failed (bits " + std::to_string(h.err) + ")");
}

inline bool scheme_ok(int s) { return s >= 0 && s <= 5; }

}  // namespace vsmcx_host

using namespace vsmcx_host;

#define API_BEGIN try {
#define API_END                                                   \
    }                                                             \
    catch (const ArgError &e) { return fail(VSMCX_ERR_ARG, e.what()); }          \
    catch (const BudgetError &e) { return fail(VSMCX_ERR_BUDGET, e.what()); }    \
    catch (const DevError &e) { return fail(e.code, e.what()); }                 \
    catch (const HipError &e) { return fail(VSMCX_ERR_HIP, e.what()); }          \
    catch (const std::exception &e) { return fail(VSMCX_ERR_INTERNAL, e.what()); } \
    return VSMCX_OK;

#define REQUIRE(cond, msg) \
    do {                   \
        if (!(cond)) throw ArgError(msg); \
    } while (0)

struct vsmcx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;  // this device's fixed array of vsmcx_resample_noalloc
    uint64_t budget = 0;
    uint64_t allocated = 0;
    std::map<void *, size_t> sizes;

    void *alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 1;
        if (budget && allocated + bytes > budget)
            throw BudgetError("device-memory budget exceeded: need " + std::to_string(bytes) +
                              " B, " + std::to_string(budget - allocated) + " B left");
        void *p = nullptr;
        VSMCX_HIP(hipSetDevice(device));
        VSMCX_HIP(hipMalloc(&p, bytes));
        allocated += bytes;
        sizes[p] = bytes;
        return p;
    }
    template <typename T>
    T *alloc_n(size_t n)
    {
        return static_cast<T *>(alloc(n * sizeof(T)));
    }
    template <typename T>
    T *alloc_zero(size_t n)
    {
        T *p = alloc_n<T>(n);
        VSMCX_HIP(hipMemsetAsync(p, 0, n * sizeof(T), stream));
        return p;
    }
    void release(void *p)
    {
        if (!p) return;
        auto it = sizes.find(p);
        if (it != sizes.end()) {
            allocated -= it->second;
            sizes.erase(it);
        }
        (void)hipFree(p);
    }
    void sync() { VSMCX_HIP(hipStreamSynchronize(stream)); }
};

namespace vsmcx {
struct ResampleBuffers;
// Fold / resample scratch for up to M particles and nuni uniforms, with the
// ancestor and cfrp lists when lists is set, and fold scratch of the uniforms'
// own when own_ufs is set (capi.cpp)
void alloc_resample(vsmcx_ctx *c, ResampleBuffers &rb, uint64_t M, uint64_t nuni, bool lists,
                    bool own_ufs = false);
void free_resample(vsmcx_ctx *c, ResampleBuffers &rb);
}  // namespace vsmcx

namespace vsmcx_host {
// The compat resample (vsmcx_resample's semantics) shared by its entry points
// (capi.cpp).  resample_args checks the arguments (ArgError) and returns false
// when there is nothing to do (M == 0).  resample_on runs it on buffers the
// caller provides: rb laid out for (M, N, copy_from && M == N), rep_out set,
// ctl one control block, w M doubles, and cf_dev either M uint64 (copy_from is
// then widened on the device and copied straight out) or NULL.
bool resample_args(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                   const vsmcx_threefry4x64_state *eng, const double *weight,
                   const uint64_t *replication);
void resample_on(vsmcx_ctx *ctx, int scheme, uint64_t M, uint64_t N,
                 vsmcx_threefry4x64_state *eng, const double *weight, uint64_t *replication,
                 uint64_t *copy_from, vsmcx::ResampleBuffers &rb, FilterCtl *ctl, double *w,
                 uint64_t *cf_dev);
// noalloc.cpp: the address of this device's fixed scratch array for
// vsmcx_resample_noalloc, looked up once per device (the device must be current)
void *noalloc_scratch(int device);
}  // namespace vsmcx_host

