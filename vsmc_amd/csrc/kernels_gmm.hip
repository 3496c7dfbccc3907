// kernels_gmm.hip -- one iteration of the Gaussian-mixture SMC sampler with
// adaptive tempering (example/gmm/src/gmm_smc.cpp under the "cess" schedule
// of example/include/smc_do.hpp), device-resident from the tempering search
// to the path-sampling record:
//
//   k_gmm_begin        bisection state (move_alpha.hpp:242-258)
//   k_gmm_weights      W_i, the normalised weights the search and the path use
//   k_gmm_cess_part    per tile: sum W b, sum W b^2 with b = exp(a_diff llh)
//   k_gmm_cess_step    one bisection step (move_alpha.hpp:259-276); the host
//                      launches the pair enough times for the search to end
//   k_gmm_reweight     lw += alpha_inc llh, log-sum-exp partials per tile
//   k_gmm_reduce       LSE, ESS, log-evidence increment, resampling decision
//   (exact-fold resampling, kernels_fold.hip)
//   k_gmm_mh           every MH sweep of every particle in one launch, the
//                      ancestor gather folded into its loads
//   k_gmm_finish       acceptance counts, path-sampling integrand and record
//
// Particle parameters are structure-of-arrays: parameter p of particle i at
// p * N + i (p < K means, then K precisions, then K weights); the log
// likelihood is FP64 per particle in both precisions.
#include "cv_model.hpp"
#include "fold_tiles.hpp"
#include "gmm_model.hpp"
#include "launch.hpp"

namespace vsmcx {

__device__ __forceinline__ void alpha2sd(double alpha, double *sd)
{
    // example/gmm/include/gmm_proposal.hpp:71-80
    alpha = alpha < 0.05 ? 0.05 : alpha;
    sd[0] = 0.15 / alpha;
    sd[1] = (1 + sqrt(1.0 / alpha)) * 0.15;
    sd[2] = (1 + sqrt(1.0 / alpha)) * 0.2;
}

// ---- initialize ------------------------------------------------------------------------
template <int K, typename Real>
__global__ __launch_bounds__(kBlock) void k_gmm_init(GmmModel g, Real *par, double *llh,
                                                     double *part)
{
    __shared__ double s_red[kBlock / kWave];
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t N = g.N;
    double v = 0;
    if (i < N) {
        GmmParticle<K, Real> p;
        gmm_init_draw<K>(g, i, p.mu, p.lam, p.w);
        const double ll = gmm_log_lik<K>(g, p.mu, p.lam, p.w);
#pragma unroll
        for (int d = 0; d < K; ++d) {
            par[(uint64_t)d * N + i] = p.mu[d];
            par[(uint64_t)(K + d) * N + i] = p.lam[d];
            par[(uint64_t)(2 * K + d) * N + i] = p.w[d];
        }
        llh[i] = ll;
        v = ll * (1.0 / (double)N);
    }
    v = block_reduce<kBlock>(v, AddOp(), s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(1024) k_gmm_init_finish(FilterCtl *ctl, GmmCtl *gc,
                                                          const double *part, uint64_t nb,
                                                          GmmRow *hist)
{
    __shared__ double s_red[1024 / kWave];
    double v = 0;
    for (uint64_t j = threadIdx.x; j < nb; j += 1024) v += part[j];
    v = block_reduce<1024>(v, AddOp(), s_red);
    if (threadIdx.x == 0) {
        gc->integrand = v;
        gc->path = 0;
        gc->zconst = 0;
        gc->alpha = 0;
        gc->alpha_inc = 0;
        gc->t = 0;
        // equal weights (core/sampler.hpp:758-771): lw = 0 implicitly
        ctl->prev_resampled = 1;
        ctl->resample = 0;
        ctl->lse = log((double)ctl->n);
        GmmRow r{};
        r.ess = (double)ctl->n;
        r.log_z = gc->llc;
        r.path_integrand = v;
        r.path_log_z = gc->llc;
        hist[0] = r;
    }
}

// ---- tempering search ------------------------------------------------------------------
__global__ void k_gmm_begin(GmmCtl *gc)
{
    if (threadIdx.x != 0) return;
    if (gc->alpha <= 0) gc->zconst = 0;
    gc->t += 1;
    gc->l = gc->alpha;
    gc->u = 1.05;
    gc->a = gc->l + 0.05;
    gc->a_old = gc->l;
    gc->done = 0;
    gc->iters = 0;
}

__global__ __launch_bounds__(kBlock) void k_gmm_weights(const FilterCtl *ctl, const double *lw,
                                                        double *W, uint64_t N)
{
    const int prev = ctl->prev_resampled;
    const double m = ctl->m, c = ctl->coeff, invN = 1.0 / (double)N;
    const uint64_t base = (uint64_t)blockIdx.x * kMoveTile;
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i < N) W[i] = prev ? invN : exp(lw[i] - m) * c;
    }
}

struct Sum2 { double a, b; };
__device__ __forceinline__ Sum2 lane_xor(const Sum2 &v, int m) { return Sum2{lane_xor(v.a, m), lane_xor(v.b, m)}; }
struct Sum2Add {
    __device__ Sum2 operator()(const Sum2 &x, const Sum2 &y) const { return Sum2{x.a + y.a, x.b + y.b}; }
};

// cess01 -> WeightSet::compute_cess (core/weight_set.hpp:411-431): the
// multipliers exp(a_diff * llh_i) are unshifted, as in the reference.
__global__ __launch_bounds__(kBlock) void k_gmm_cess_part(const GmmCtl *gc, const double *W,
                                                          const double *llh, uint64_t N,
                                                          double *part)
{
    if (gc->done) return;
    __shared__ Sum2 s_r[kBlock / kWave];
    const double a_diff = gc->a - gc->a_old;
    const uint64_t base = (uint64_t)blockIdx.x * kMoveTile;
    Sum2 v{0, 0};
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i < N) {
            const double b = exp(a_diff * llh[i]);
            const double wb = W[i] * b;
            v.a += wb;
            v.b += wb * b;
        }
    }
    v = block_reduce<kBlock>(v, Sum2Add(), s_r);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v.a;
        part[2 * blockIdx.x + 1] = v.b;
    }
}

__global__ void __launch_bounds__(1024) k_gmm_cess_step(GmmCtl *gc, const double *part,
                                                        uint64_t ntiles)
{
    if (gc->done) return;
    __shared__ Sum2 s_r[1024 / kWave];
    Sum2 v{0, 0};
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) {
        v.a += part[2 * j];
        v.b += part[2 * j + 1];
    }
    v = block_reduce<1024>(v, Sum2Add(), s_r);
    if (threadIdx.x != 0) return;
    const double cess = v.a * v.a / v.b;
    gc->iters += 1;
    double l = gc->l, u = gc->u, a = gc->a;
    if (cess < gc->target) {
        u = a;
        a = 0.5 * (a + l);
    } else {
        l = a;
        a = 0.5 * (a + u);
    }
    gc->l = l;
    gc->u = u;
    gc->a = a;
    if (fabs(u - l) > gc->eps && l <= 1) return;
    // gmm_state::alpha (example/gmm/include/gmm_state.hpp:163-172)
    a = a < 1 ? a : 1;
    a = a > 0 ? a : 0;
    gc->alpha_inc = a > 0 ? a - gc->alpha : 0;
    gc->alpha = a;
    alpha2sd(a, gc->sd);
    gc->done = 1;
}

// ---- reweight and normalise ------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_gmm_reweight(const FilterCtl *ctl, const GmmCtl *gc,
                                                         const double *llh, double *lw,
                                                         uint64_t N, double *part)
{
    __shared__ double s_red[kBlock / kWave];
    __shared__ Sum2 s_r[kBlock / kWave];
    const int prev = ctl->prev_resampled;
    const double ainc = gc->alpha_inc;
    const uint64_t base = (uint64_t)blockIdx.x * kMoveTile;
    double m = -INFINITY, se = 0, se2 = 0;
#pragma unroll
    for (int k = 0; k < kMoveItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        if (i < N) {
            const double v = (prev ? 0.0 : lw[i]) + llh[i] * ainc;
            lw[i] = v;
            lse_add(m, se, se2, v);
        }
    }
    const double M = block_reduce<kBlock>(m, MaxOp(), s_red);
    const double sc = (m == -INFINITY) ? 0.0 : exp(m - M);
    Sum2 v{se * sc, se2 * sc * sc};
    v = block_reduce<kBlock>(v, Sum2Add(), s_r);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = M;
        part[3 * blockIdx.x + 1] = v.a;
        part[3 * blockIdx.x + 2] = v.b;
    }
}

__global__ void __launch_bounds__(1024) k_gmm_reduce(FilterCtl *ctl, GmmCtl *gc,
                                                     const double *part, uint64_t ntiles,
                                                     GmmRow *hist, int scheme)
{
    __shared__ double s_red[1024 / kWave];
    __shared__ Sum2 s_r[1024 / kWave];
    double m = -INFINITY;
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) m = fmax(m, part[3 * j]);
    const double M = block_reduce<1024>(m, MaxOp(), s_red);
    Sum2 acc{0, 0};
    for (uint64_t j = threadIdx.x; j < ntiles; j += 1024) {
        const double *q = part + 3 * j;
        const double sc = q[0] == -INFINITY ? 0.0 : exp(q[0] - M);
        acc.a += q[1] * sc;
        acc.b += q[2] * sc * sc;
    }
    acc = block_reduce<1024>(acc, Sum2Add(), s_r);
    if (threadIdx.x != 0) return;
    if (!gc->done) atomicOr(&ctl->err, (int)kErrSearch);  // the search did not finish
    const double n = (double)ctl->n;
    const double se = acc.a, se2 = acc.b;
    const double lse_old = ctl->prev_resampled ? log(n) : ctl->lse;
    const double lse = M + log(se);
    // log sum_i W_i exp(inc_i) = LSE(lw + inc) - LSE(lw)  (move_smc.hpp:64-71)
    gc->zconst += lse - lse_old;
    const double ess = se * se / se2;
    int res = ess < ctl->threshold * n;
    if (!(se > 0) || !isfinite(lse) || !isfinite(ess)) {
        atomicOr(&ctl->err, (int)kErrDegenerate);
        res = 0;
    }
    ctl->m = M;
    ctl->coeff = 1 / se;
    ctl->lse = lse;
    ctl->ess = ess;
    ctl->resample = res;
    ctl->t = gc->t;
    ctl->prev_resampled = 0;
    ctl->r_int = 0;
    ctl->draws = 0;
    if (res) resample_prologue(ctl, scheme);
    GmmRow &r = hist[gc->t];
    r.alpha = gc->alpha;
    r.alpha_inc = gc->alpha_inc;
    r.ess = ess;
    r.log_z = gc->zconst + gc->llc;
    r.resampled = res;
    r.bisections = gc->iters;
}

// ---- MH sweeps ---------------------------------------------------------------------------
template <int K, typename Real>
__global__ __launch_bounds__(kBlock) void k_gmm_mh(GmmModel g, const FilterCtl *ctl,
                                                   const GmmCtl *gc, const Real *pin, Real *pout,
                                                   const double *llh_in, double *llh_out,
                                                   const uint32_t *anc, const double *lw,
                                                   double *ppart, uint32_t *apart)
{
    __shared__ double s_red[kBlock / kWave];
    __shared__ int s_acc[kBlock / kWave];
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t N = g.N;
    const double alpha = gc->alpha;
    const uint32_t iter = (uint32_t)gc->t;
    double sd[3] = {gc->sd[0], gc->sd[1], gc->sd[2]};
    int acc[3] = {0, 0, 0};
    double pv = 0;
    if (i < N) {
        uint64_t src = ctl->resample ? (uint64_t)anc[i] : i;
        src = src < N ? src : i;
        GmmParticle<K, Real> p;
#pragma unroll
        for (int d = 0; d < K; ++d) {
            p.mu[d] = pin[(uint64_t)d * N + src];
            p.lam[d] = pin[(uint64_t)(K + d) * N + src];
            p.w[d] = pin[(uint64_t)(2 * K + d) * N + src];
        }
        p.llh = llh_in[src];
        p.lp = gmm_log_prior<K>(g, p.mu, p.lam);
#pragma unroll 1
        for (uint32_t s = 0; s < g.sweeps; ++s) {
            acc[0] += gmm_mh_move<K>(g, p, 0, i, iter, s, alpha, sd);
            acc[1] += gmm_mh_move<K>(g, p, 1, i, iter, s, alpha, sd);
            acc[2] += gmm_mh_move<K>(g, p, 2, i, iter, s, alpha, sd);
        }
#pragma unroll
        for (int d = 0; d < K; ++d) {
            pout[(uint64_t)d * N + i] = p.mu[d];
            pout[(uint64_t)(K + d) * N + i] = p.lam[d];
            pout[(uint64_t)(2 * K + d) * N + i] = p.w[d];
        }
        llh_out[i] = p.llh;
        // path-sampling integrand sum_i W_i llh_i at MonitorMCMC (after the
        // resample the weights are equal)
        pv = (ctl->resample ? 1.0 / (double)N : exp(lw[i] - ctl->m) * ctl->coeff) * p.llh;
    }
    pv = block_reduce<kBlock>(pv, AddOp(), s_red);
    int a0 = block_reduce<kBlock>(acc[0], AddOp(), s_acc);
    int a1 = block_reduce<kBlock>(acc[1], AddOp(), s_acc);
    int a2 = block_reduce<kBlock>(acc[2], AddOp(), s_acc);
    if (threadIdx.x == 0) {
        ppart[blockIdx.x] = pv;
        apart[3 * blockIdx.x] = (uint32_t)a0;
        apart[3 * blockIdx.x + 1] = (uint32_t)a1;
        apart[3 * blockIdx.x + 2] = (uint32_t)a2;
    }
}

__global__ void __launch_bounds__(1024) k_gmm_finish(GmmCtl *gc, const double *ppart,
                                                     const uint32_t *apart, uint64_t nb,
                                                     GmmRow *hist)
{
    __shared__ double s_red[1024 / kWave];
    __shared__ uint64_t s_u[1024 / kWave];
    double v = 0;
    uint64_t a[3] = {0, 0, 0};
    for (uint64_t j = threadIdx.x; j < nb; j += 1024) {
        v += ppart[j];
        a[0] += apart[3 * j];
        a[1] += apart[3 * j + 1];
        a[2] += apart[3 * j + 2];
    }
    v = block_reduce<1024>(v, AddOp(), s_red);
    uint64_t t0 = block_reduce<1024>(a[0], AddOp(), s_u);
    uint64_t t1 = block_reduce<1024>(a[1], AddOp(), s_u);
    uint64_t t2 = block_reduce<1024>(a[2], AddOp(), s_u);
    if (threadIdx.x != 0) return;
    GmmRow &r = hist[gc->t];
    const double a_prev = hist[gc->t - 1].alpha;
    // core/path.hpp:229-239: trapezoid on the alpha grid
    gc->path += 0.5 * (gc->alpha - a_prev) * (v + gc->integrand);
    gc->integrand = v;
    r.path_integrand = v;
    r.path_log_z = gc->path + gc->llc;
    r.accept[0] = t0;
    r.accept[1] = t1;
    r.accept[2] = t2;
}

// ---- launchers ---------------------------------------------------------------------------
template <typename Real>
static void gmm_init_k(int K, const GmmModel &g, Real *par, double *llh, double *part,
                       hipStream_t st)
{
    const dim3 grid((unsigned)((g.N + kBlock - 1) / kBlock)), b(kBlock);
    switch (K) {
    case 1: hipLaunchKernelGGL((k_gmm_init<1, Real>), grid, b, 0, st, g, par, llh, part); break;
    case 2: hipLaunchKernelGGL((k_gmm_init<2, Real>), grid, b, 0, st, g, par, llh, part); break;
    case 3: hipLaunchKernelGGL((k_gmm_init<3, Real>), grid, b, 0, st, g, par, llh, part); break;
    case 4: hipLaunchKernelGGL((k_gmm_init<4, Real>), grid, b, 0, st, g, par, llh, part); break;
    case 5: hipLaunchKernelGGL((k_gmm_init<5, Real>), grid, b, 0, st, g, par, llh, part); break;
    default: hipLaunchKernelGGL((k_gmm_init<6, Real>), grid, b, 0, st, g, par, llh, part); break;
    }
}

void launch_gmm_init(bool fp64, int K, const GmmModel &g, FilterCtl *ctl, GmmCtl *gc, void *par,
                     double *llh, double *part, GmmRow *hist, hipStream_t st)
{
    if (fp64)
        gmm_init_k<double>(K, g, (double *)par, llh, part, st);
    else
        gmm_init_k<float>(K, g, (float *)par, llh, part, st);
    const uint64_t nb = (g.N + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gmm_init_finish, dim3(1), dim3(1024), 0, st, ctl, gc, part, nb, hist);
}

void launch_gmm_temper(FilterCtl *ctl, GmmCtl *gc, const double *llh, double *lw, double *W,
                       uint64_t N, uint32_t passes, double *part, GmmRow *hist, int scheme,
                       hipStream_t st)
{
    const uint64_t tiles = (N + kMoveTile - 1) / kMoveTile;
    const dim3 gt((unsigned)tiles), b(kBlock);
    hipLaunchKernelGGL(k_gmm_begin, dim3(1), dim3(64), 0, st, gc);
    hipLaunchKernelGGL(k_gmm_weights, gt, b, 0, st, ctl, lw, W, N);
    for (uint32_t k = 0; k < passes; ++k) {
        hipLaunchKernelGGL(k_gmm_cess_part, gt, b, 0, st, gc, W, llh, N, part);
        hipLaunchKernelGGL(k_gmm_cess_step, dim3(1), dim3(1024), 0, st, gc, part, tiles);
    }
    hipLaunchKernelGGL(k_gmm_reweight, gt, b, 0, st, ctl, gc, llh, lw, N, part);
    hipLaunchKernelGGL(k_gmm_reduce, dim3(1), dim3(1024), 0, st, ctl, gc, part, tiles, hist,
                       scheme);
}

template <typename Real>
static void gmm_mh_k(int K, const GmmModel &g, const FilterCtl *ctl, const GmmCtl *gc,
                     const Real *pin, Real *pout, const double *li, double *lo,
                     const uint32_t *anc, const double *lw, double *pp, uint32_t *ap,
                     hipStream_t st)
{
    const dim3 grid((unsigned)((g.N + kBlock - 1) / kBlock)), b(kBlock);
#define VSMCX_GMM_MH(KK)                                                                       \
    hipLaunchKernelGGL((k_gmm_mh<KK, Real>), grid, b, 0, st, g, ctl, gc, pin, pout, li, lo, anc, \
                       lw, pp, ap)
    switch (K) {
    case 1: VSMCX_GMM_MH(1); break;
    case 2: VSMCX_GMM_MH(2); break;
    case 3: VSMCX_GMM_MH(3); break;
    case 4: VSMCX_GMM_MH(4); break;
    case 5: VSMCX_GMM_MH(5); break;
    default: VSMCX_GMM_MH(6); break;
    }
#undef VSMCX_GMM_MH
}

void launch_gmm_mh(bool fp64, int K, const GmmModel &g, const FilterCtl *ctl, GmmCtl *gc,
                   const void *pin, void *pout, const double *llh_in, double *llh_out,
                   const uint32_t *anc, const double *lw, double *ppart, uint32_t *apart,
                   GmmRow *hist, hipStream_t st)
{
    if (fp64)
        gmm_mh_k<double>(K, g, ctl, gc, (const double *)pin, (double *)pout, llh_in, llh_out, anc,
                         lw, ppart, apart, st);
    else
        gmm_mh_k<float>(K, g, ctl, gc, (const float *)pin, (float *)pout, llh_in, llh_out, anc,
                        lw, ppart, apart, st);
    const uint64_t nb = (g.N + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gmm_finish, dim3(1), dim3(1024), 0, st, gc, ppart, apart, nb, hist);
}

}  // namespace vsmcx
