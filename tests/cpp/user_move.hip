// A move written as a user kernel: the OpenCL pf example's cv_init and
// cv_move (example/pf/pf_cl.cl:46-122) in HIP, with fp_type double and
// NORMAL01_4x32 = normal01_4x32_53 from include/vsmcx_device.hpp, over a
// vsmcx_state and the device weight set (vsmcx_weights_*).  Each slot keeps
// its engine counter across iterations (the example's counter buffer, which
// is user memory here too).  Nothing crosses PCIe inside the loop.
//
// Usage: user_move N scheme seed resample_seed threshold < obs   (T lines "x y")
// Output: T lines "t ess log_z resampled draws", then N lines of final state.
//
// Test infrastructure: built by tests/cpp/Makefile, run on a GPU by
// tests/test_gpu_user_move.py.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vsmcx.h"
#include "vsmcx_device.hpp"

using namespace vsmcx;

struct cv {
    double x_pos, y_pos, x_vel, y_vel;
};

__device__ double log_likelihood(const cv &sp, double x_obs, double y_obs)
{
    const double scale = 10, nu = 10;
    double llh_x = scale * (sp.x_pos - x_obs);
    double llh_y = scale * (sp.y_pos - y_obs);
    llh_x = log1p(llh_x * llh_x / nu);
    llh_y = log1p(llh_y * llh_y / nu);
    return -0.5 * (nu + 1) * (llh_x + llh_y);
}

__global__ void cv_init(cv *state, double *log_weight, double x_obs, double y_obs, u32x4 *counter,
                        uint32_t seed, uint64_t size)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= size) return;
    const double sd_pos0 = 2, sd_vel0 = 1;
    cv sp;
    cburng4x32_rng_t rng;
    cburng4x32_init(&rng, (uint32_t)(seed + i));
    normal01_4x32_53 rnorm;
    normal01_4x32_53_init(&rnorm, &rng);
    sp.x_pos = normal01_4x32_53_rand(&rnorm, &rng) * sd_pos0;
    sp.y_pos = normal01_4x32_53_rand(&rnorm, &rng) * sd_pos0;
    sp.x_vel = normal01_4x32_53_rand(&rnorm, &rng) * sd_vel0;
    sp.y_vel = normal01_4x32_53_rand(&rnorm, &rng) * sd_vel0;
    state[i] = sp;
    log_weight[i] = log_likelihood(sp, x_obs, y_obs);
    counter[i] = rng.ctr;
}

__global__ void cv_move(uint64_t iter, cv *state, double *inc_weight, double x_obs, double y_obs,
                        u32x4 *counter, uint32_t seed, uint64_t size)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= size) return;
    (void)iter;
    const double var_pos = 0.02, var_vel = 0.001;
    const double sd_pos = sqrt(var_pos), sd_vel = sqrt(var_vel);
    const double delta = 0.1;
    cv sp = state[i];
    cburng4x32_rng_t rng;
    cburng4x32_init(&rng, (uint32_t)(seed + i));
    rng.ctr = counter[i];
    normal01_4x32_53 rnorm;
    normal01_4x32_53_init(&rnorm, &rng);
    sp.x_pos += normal01_4x32_53_rand(&rnorm, &rng) * sd_pos + delta * sp.x_vel;
    sp.y_pos += normal01_4x32_53_rand(&rnorm, &rng) * sd_pos + delta * sp.y_vel;
    sp.x_vel += normal01_4x32_53_rand(&rnorm, &rng) * sd_vel;
    sp.y_vel += normal01_4x32_53_rand(&rnorm, &rng) * sd_vel;
    state[i] = sp;
    inc_weight[i] = log_likelihood(sp, x_obs, y_obs);
    counter[i] = rng.ctr;
}

static void ok(int rc, const char *what)
{
    if (rc != VSMCX_OK) {
        std::fprintf(stderr, "error: %s: %s\n", what, vsmcx_last_error());
        std::exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: user_move N scheme seed resample_seed threshold < obs\n");
        return 2;
    }
    const uint64_t N = std::strtoull(argv[1], nullptr, 10);
    const int scheme = std::atoi(argv[2]);
    const uint32_t seed = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const uint64_t rseed = std::strtoull(argv[4], nullptr, 10);
    const double threshold = std::atof(argv[5]);
    std::vector<double> obs;
    double x, y;
    while (std::scanf("%lf %lf", &x, &y) == 2) {
        obs.push_back(x);
        obs.push_back(y);
    }
    const uint64_t T = obs.size() / 2;
    if (T == 0 || N == 0) {
        std::fprintf(stderr, "no observations or no particles\n");
        return 2;
    }

    vsmcx_ctx *ctx = nullptr;
    ok(vsmcx_ctx_create(0, 0, &ctx), "context");
    void *sp = nullptr;
    ok(vsmcx_ctx_stream(ctx, &sp), "stream");
    hipStream_t stream = (hipStream_t)sp;
    vsmcx_state *state = nullptr;
    ok(vsmcx_state_create(ctx, N, 4, VSMCX_FP64_FUSED, &state), "state");
    void *rows = nullptr;
    ok(vsmcx_state_data(state, &rows), "state data");
    vsmcx_weights *w = nullptr;
    ok(vsmcx_weights_create(ctx, N, scheme, threshold, rseed, T, &w), "weights");
    double *inc = nullptr;
    ok(vsmcx_weights_buffer(w, &inc), "increment buffer");
    u32x4 *counter = nullptr;
    if (hipMalloc(&counter, N * sizeof(u32x4)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }

    const unsigned block = 256, grid = (unsigned)((N + block - 1) / block);
    hipLaunchKernelGGL(cv_init, dim3(grid), dim3(block), 0, stream, (cv *)rows, inc, obs[0], obs[1],
                       counter, seed, N);
    ok(vsmcx_weights_set_log(w, inc), "set_log");
    ok(vsmcx_weights_resample(w, state), "resample");
    for (uint64_t t = 1; t < T; ++t) {
        hipLaunchKernelGGL(cv_move, dim3(grid), dim3(block), 0, stream, t, (cv *)rows, inc,
                           obs[2 * t], obs[2 * t + 1], counter, seed, N);
        ok(vsmcx_weights_add_log(w, inc), "add_log");
        ok(vsmcx_weights_resample(w, state), "resample");
    }
    if (hipGetLastError() != hipSuccess) {
        std::fprintf(stderr, "kernel launch failed\n");
        return 1;
    }

    std::vector<vsmcx_step_stats> h(T);
    ok(vsmcx_weights_history(w, T, h.data()), "history");
    for (uint64_t t = 0; t < T; ++t)
        std::printf("%llu %.17g %.17g %d %llu\n", (unsigned long long)t, h[t].ess, h[t].log_z,
                    h[t].resampled, (unsigned long long)h[t].draws);
    std::vector<double> st(4 * N);
    ok(vsmcx_state_read(state, st.data()), "state read");
    for (uint64_t i = 0; i < N; ++i)
        std::printf("%.17g %.17g %.17g %.17g\n", st[4 * i], st[4 * i + 1], st[4 * i + 2],
                    st[4 * i + 3]);
    (void)hipFree(counter);
    vsmcx_weights_destroy(w);
    vsmcx_state_destroy(state);
    vsmcx_ctx_destroy(ctx);
    return 0;
}
