// ref_harness.cpp -- extern "C" shims over the unmodified vSMC headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// headers where they lie (-I <vsmc>/include) into oracle/_ref/libvsmcref.so.
// It is used (a) to write the golden vectors under tests/golden/ and (b) to
// cross-check the C restatement in oracle/vsmc_oracle.c.  The product never
// loads it.  No reference source is copied here: this file only instantiates
// vSMC's own templates.
#include <vsmc/core/particle.hpp>
#include <vsmc/core/state_matrix.hpp>
#include <vsmc/core/weight_set.hpp>
#include <vsmc/resample/common.hpp>
#include <vsmc/resample/multinomial.hpp>
#include <vsmc/resample/residual.hpp>
#include <vsmc/resample/residual_stratified.hpp>
#include <vsmc/resample/residual_systematic.hpp>
#include <vsmc/resample/stratified.hpp>
#include <vsmc/resample/systematic.hpp>
#include <vsmc/rng/u01.h>
#include <vsmc/rng/philox.hpp>
#include <vsmc/rng/threefry.hpp>
#include <vsmc/rng/u01.hpp>
#include <vsmc/smp/backend_seq.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

// --- the gmm example's tempering search, driven on a harness value type ------
// move_alpha.hpp (example/include) needs only vsmc::Particle<T> and the value
// type's state(i, 0).log_likelihood(), state(0, 0).alpha() and alpha(a); the
// gmm example's own gmm_state cannot be included here (its header pulls in
// example/include/common.hpp, which includes core/sampler.hpp, which does not
// compile with this toolchain).  The harness type below exposes exactly those
// members; alpha(a) is the clamp and alpha_inc of gmm_state::alpha
// (example/gmm/include/gmm_state.hpp:163-172).  Its weight set counts the
// compute_cess calls (a virtual of vsmc::WeightSet), i.e. the bisection steps.
#include <move_alpha.hpp>  // example/include, on the include path (oracle/Makefile)

struct HarnessParam {
    double ll = 0, a = 0, a_inc = 0;
    double &log_likelihood() { return ll; }
    double log_likelihood() const { return ll; }
    double &alpha() { return a; }
    double alpha() const { return a; }
    double &alpha_inc() { return a_inc; }
};

class CountingWeightSet : public vsmc::WeightSet
{
    public :
    explicit CountingWeightSet(size_type N) : vsmc::WeightSet(N), calls(0) {}
    mutable std::size_t calls;
    double compute_cess(const double *first, bool use_log) const
    {
        ++calls;
        return vsmc::WeightSet::compute_cess(first, use_log);
    }
};

class HarnessGmmState : public vsmc::StateSEQ<vsmc::StateMatrix<vsmc::RowMajor, 1, HarnessParam> >
{
    public :
    typedef CountingWeightSet weight_set_type;
    explicit HarnessGmmState(size_type N)
        : vsmc::StateSEQ<vsmc::StateMatrix<vsmc::RowMajor, 1, HarnessParam> >(N) {}
    void alpha(double a)
    {
        a = a < 1 ? a : 1;
        a = a > 0 ? a : 0;
        double a_inc = a > 0 ? a - this->state(0, 0).alpha() : 0;
        for (size_type i = 0; i != this->size(); ++i) {
            this->state(i, 0).alpha() = a;
            this->state(i, 0).alpha_inc() = a_inc;
        }
    }
};

typedef vsmc::Philox4x32 Ph32;
typedef vsmc::Threefry4x64 Tf64;
typedef vsmc::Threefry4x32 Tf32;

template <typename Eng, typename T>
static void draw_n(Eng &eng, std::size_t n, T *out)
{
    for (std::size_t i = 0; i != n; ++i) out[i] = static_cast<T>(eng());
}

extern "C" {

// --- engine blocks: the stateless operator()(ctr) of each engine -----------
void ref_philox4x32_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    Ph32 eng;
    Ph32::key_type k;
    k[0] = key[0]; k[1] = key[1];
    eng.key(k);
    Ph32::ctr_type c;
    for (int j = 0; j < 4; ++j) c[j] = ctr[j];
    Ph32::buffer_type b = eng(c);
    for (int j = 0; j < 4; ++j) out[j] = b[j];
}

void ref_threefry4x64_block(const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4])
{
    Tf64::key_type k;
    for (int j = 0; j < 4; ++j) k[j] = key[j];
    Tf64 eng(k);
    Tf64::ctr_type c;
    for (int j = 0; j < 4; ++j) c[j] = ctr[j];
    Tf64::buffer_type b = eng(c);
    for (int j = 0; j < 4; ++j) out[j] = b[j];
}

void ref_threefry4x32_block(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4])
{
    Tf32::key_type k;
    for (int j = 0; j < 4; ++j) k[j] = key[j];
    Tf32 eng(k);
    Tf32::ctr_type c;
    for (int j = 0; j < 4; ++j) c[j] = ctr[j];
    Tf32::buffer_type b = eng(c);
    for (int j = 0; j < 4; ++j) out[j] = b[j];
}

// --- engine streams: seed(s), optional ctr set, discard, n draws ----------
void ref_philox4x32_stream(uint32_t seed, const uint32_t *ctr, uint64_t skip, size_t n,
                           uint32_t *out)
{
    Ph32 eng(seed);
    if (ctr) {
        Ph32::ctr_type c;
        for (int j = 0; j < 4; ++j) c[j] = ctr[j];
        eng.ctr(c);
    }
    if (skip) eng.discard(static_cast<Ph32::result_type>(skip));
    draw_n(eng, n, out);
}

void ref_threefry4x64_stream(uint64_t seed, const uint64_t *ctr, uint64_t skip, size_t n,
                             uint64_t *out)
{
    Tf64 eng(seed);
    if (ctr) {
        Tf64::ctr_type c;
        for (int j = 0; j < 4; ++j) c[j] = ctr[j];
        eng.ctr(c);
    }
    if (skip) eng.discard(static_cast<Tf64::result_type>(skip));
    draw_n(eng, n, out);
}

void ref_threefry4x32_stream(uint32_t seed, size_t n, uint32_t *out)
{
    Tf32 eng(seed);
    draw_n(eng, n, out);
}

// --- std::uniform_real_distribution<double> through vSMC's cxx11 alias ----
void ref_canonical_tf64(uint64_t seed, size_t n, double *out)
{
    Tf64 eng(seed);
    vsmc::cxx11::uniform_real_distribution<double> runif(0, 1);
    for (size_t i = 0; i < n; ++i) out[i] = runif(eng);
}

void ref_canonical_ph32(uint32_t seed, size_t n, double *out)
{
    Ph32 eng(seed);
    vsmc::cxx11::uniform_real_distribution<double> runif(0, 1);
    for (size_t i = 0; i < n; ++i) out[i] = runif(eng);
}

// --- U01Sequence{Sorted,Stratified,Systematic}: all N values, then the next
//     4 engine draws (pins how far the engine advanced) -----------------------
void ref_u01seq(int kind, size_t N, uint64_t seed, double *u, uint64_t next4[4])
{
    Tf64 eng(seed);
    if (kind == 0) {
        vsmc::U01SequenceSorted<Tf64> s(N, eng);
        for (size_t n = 0; n < N; ++n) u[n] = s[n];
    } else if (kind == 1) {
        vsmc::U01SequenceStratified<Tf64> s(N, eng);
        for (size_t n = 0; n < N; ++n) u[n] = s[n];
    } else {
        vsmc::U01SequenceSystematic<Tf64> s(N, eng);
        for (size_t n = 0; n < N; ++n) u[n] = s[n];
    }
    draw_n(eng, 4, next4);
}

// --- Resample<Scheme>(M=N, N, rng, w, rep) then cfrp_trans; next 4 draws ---
void ref_resample(int scheme, size_t N, uint64_t seed, const double *w, uint64_t *rep,
                  uint64_t *copy_from, uint64_t next4[4])
{
    Tf64 eng(seed);
    std::vector<std::size_t> r(N), cf(N);
    switch (scheme) {
    case 0: { vsmc::ResampleType<vsmc::Multinomial>::type op; op(N, N, eng, w, &r[0]); break; }
    case 1: { vsmc::ResampleType<vsmc::Residual>::type op; op(N, N, eng, w, &r[0]); break; }
    case 2: { vsmc::ResampleType<vsmc::Stratified>::type op; op(N, N, eng, w, &r[0]); break; }
    case 3: { vsmc::ResampleType<vsmc::Systematic>::type op; op(N, N, eng, w, &r[0]); break; }
    case 4: {
        vsmc::ResampleType<vsmc::ResidualStratified>::type op;
        op(N, N, eng, w, &r[0]);
        break;
    }
    default: {
        vsmc::ResampleType<vsmc::ResidualSystematic>::type op;
        op(N, N, eng, w, &r[0]);
        break;
    }
    }
    vsmc::internal::cfrp_trans(N, N, &r[0], &cf[0]);
    for (size_t i = 0; i < N; ++i) {
        rep[i] = r[i];
        copy_from[i] = cf[i];
    }
    draw_n(eng, 4, next4);
}

// --- internal::cfrp_trans on a given replication vector -------------------
void ref_cfrp_trans(size_t M, size_t N, const uint64_t *rep, uint64_t *copy_from)
{
    std::vector<std::size_t> r(rep, rep + M), cf(N);
    vsmc::internal::cfrp_trans(M, N, &r[0], &cf[0]);
    for (size_t i = 0; i < N; ++i) copy_from[i] = cf[i];
}

// --- WeightSet::set_log_weight / add_log_weight ----------------------------
double ref_weightset(size_t N, const double *lw0, const double *inc, double *lw_out,
                     double *w_out)
{
    vsmc::WeightSet ws(N);
    ws.set_log_weight(lw0);
    if (inc) ws.add_log_weight(inc);
    ws.read_log_weight(lw_out);
    ws.read_weight(w_out);
    return ws.ess();
}

// --- WeightSet::cess (core/weight_set.hpp:161-186, 411-431) -----------------
// weights from set_log_weight(lw) (read back into w_out); mul are the
// multipliers (use_log 0) or their logarithms (use_log 1)
double ref_cess(size_t N, const double *lw, const double *mul, int use_log, double *w_out)
{
    vsmc::WeightSet ws(N);
    ws.set_log_weight(lw);
    if (w_out) ws.read_weight(w_out);
    return ws.cess(mul, use_log != 0);
}

// --- alpha_ess<T, cess01<T>>::alpha_iter (example/include/move_alpha.hpp:
// 234-279) on a Particle<HarnessGmmState> with log-weights lw, per-particle
// log-likelihoods llh and current temperature alpha; returns the new alpha,
// alpha_inc, and the number of CESS evaluations; w_out: the weights it used
void ref_alpha_ess(size_t N, const double *lw, const double *llh, double alpha, double drop,
                   double *alpha_out, double *alpha_inc_out, uint64_t *cess_calls, double *w_out)
{
    vsmc::Particle<HarnessGmmState> particle(N);
    particle.weight_set().set_log_weight(lw);
    if (w_out) particle.weight_set().read_weight(w_out);
    for (size_t i = 0; i < N; ++i) {
        particle.value().state(i, 0).log_likelihood() = llh[i];
        particle.value().state(i, 0).alpha() = alpha;
    }
    particle.weight_set().calls = 0;
    alpha_ess<HarnessGmmState, cess01<HarnessGmmState> > search(drop);
    search.alpha_iter(1, particle);
    *alpha_out = particle.value().state(0, 0).alpha();
    *alpha_inc_out = particle.value().state(0, 0).alpha_inc();
    *cess_calls = particle.weight_set().calls;
}

// --- CV particle filter driven through vsmc::Particle<T> -------------------
// Replays Sampler::do_init / do_iter (core/sampler.hpp:773-834) on a
// Particle<StateMatrix<RowMajor, 4, double>>; the model arithmetic is the one
// the workload states (normals from a Philox4x32 engine whose counter is set
// to {t, id_lo, id_hi, 0} before one block is drawn; Box-Muller on
// u01_open_closed_32_53; Student-t likelihood with log1p).
struct ref_cv_param {
    double dt, sd_pos, sd_vel, prior_sd_pos, prior_sd_vel;
    double obs_scale, nu, llh_const;
};
struct ref_step_record {
    double ess, log_z, mean_x, mean_y;
    int resampled;
    uint64_t draws;
};

typedef vsmc::StateMatrix<vsmc::RowMajor, 4, double> CvState;

static void cv_normals(uint64_t seed, uint64_t id, uint64_t t, double z[4])
{
    Ph32::key_type k;
    k[0] = static_cast<uint32_t>(seed);
    k[1] = static_cast<uint32_t>(seed >> 32);
    Ph32 eng(k);
    Ph32::ctr_type c;
    c[0] = static_cast<uint32_t>(t);
    c[1] = static_cast<uint32_t>(id);
    c[2] = static_cast<uint32_t>(id >> 32);
    c[3] = 0;
    eng.ctr(c);
    uint32_t w[4];
    for (int j = 0; j < 4; ++j) w[j] = eng();
    const double two_pi = 6.2831853071795865;
    for (int j = 0; j < 2; ++j) {
        double u1 = ::u01_open_closed_32_53(w[2 * j]);
        double u2 = ::u01_open_closed_32_53(w[2 * j + 1]);
        double r = std::sqrt(-2 * std::log(u1));
        z[2 * j] = r * std::cos(two_pi * u2);
        z[2 * j + 1] = r * std::sin(two_pi * u2);
    }
}

static double cv_llh(const ref_cv_param &p, double x, double y, double ox, double oy)
{
    double dx = (x - ox) / p.obs_scale;
    double dy = (y - oy) / p.obs_scale;
    double lx = std::log1p(dx * dx / p.nu);
    double ly = std::log1p(dy * dy / p.nu);
    return 2 * p.llh_const - 0.5 * (p.nu + 1) * (lx + ly);
}

int ref_cv_run(const ref_cv_param *p, size_t N, size_t T, const double *obs, uint64_t seed,
               uint64_t resample_seed, int scheme, double threshold, double *state_out,
               double *lw_out, ref_step_record *rec)
{
    vsmc::Particle<CvState> particle(N);
    particle.resample_rng().seed(resample_seed);
    vsmc::ResampleScheme schemes[6] = {vsmc::Multinomial, vsmc::Residual, vsmc::Stratified,
                                       vsmc::Systematic, vsmc::ResidualStratified,
                                       vsmc::ResidualSystematic};
    std::function<void(std::size_t, std::size_t, Tf64 &, const double *, std::size_t *)> op;
    switch (schemes[scheme]) {
    case vsmc::Multinomial: op = vsmc::ResampleType<vsmc::Multinomial>::type(); break;
    case vsmc::Residual: op = vsmc::ResampleType<vsmc::Residual>::type(); break;
    case vsmc::Stratified: op = vsmc::ResampleType<vsmc::Stratified>::type(); break;
    case vsmc::Systematic: op = vsmc::ResampleType<vsmc::Systematic>::type(); break;
    case vsmc::ResidualStratified:
        op = vsmc::ResampleType<vsmc::ResidualStratified>::type();
        break;
    default: op = vsmc::ResampleType<vsmc::ResidualSystematic>::type(); break;
    }
    // count engine draws by wrapping the op around a draw-counting copy
    particle.weight_set().set_equal_weight();  // do_reset
    std::vector<double> inc(N), w(N);
    double log_z = 0;
    for (size_t t = 0; t < T; ++t) {
        const double ox = obs[2 * t], oy = obs[2 * t + 1];
        CvState &v = particle.value();
        for (size_t i = 0; i < N; ++i) {
            double z[4];
            cv_normals(seed, i, t, z);
            if (t == 0) {
                v.state(i, 0) = p->prior_sd_pos * z[0];
                v.state(i, 1) = p->prior_sd_pos * z[1];
                v.state(i, 2) = p->prior_sd_vel * z[2];
                v.state(i, 3) = p->prior_sd_vel * z[3];
            } else {
                v.state(i, 0) += p->sd_pos * z[0] + p->dt * v.state(i, 2);
                v.state(i, 1) += p->sd_pos * z[1] + p->dt * v.state(i, 3);
                v.state(i, 2) += p->sd_vel * z[2];
                v.state(i, 3) += p->sd_vel * z[3];
            }
            inc[i] = cv_llh(*p, v.state(i, 0), v.state(i, 1), ox, oy);
        }
        // move_smc.hpp:64-71 style increment, with a max shift
        particle.weight_set().read_weight(&w[0]);
        double imax = inc[0];
        for (size_t i = 0; i < N; ++i)
            if (imax < inc[i]) imax = inc[i];
        double sum = 0;
        for (size_t i = 0; i < N; ++i) sum += w[i] * std::exp(inc[i] - imax);
        log_z += std::log(sum) + imax;
        if (t == 0)
            particle.weight_set().set_log_weight(&inc[0]);
        else
            particle.weight_set().add_log_weight(&inc[0]);
        double ess = particle.weight_set().ess();
        // draws: compare engine before/after via a copy and stepping
        Tf64 before = particle.resample_rng();
        bool resampled = particle.resample(op, threshold);
        uint64_t draws = 0;
        if (resampled) {
            Tf64 probe = before;
            while (!(probe == particle.resample_rng()) && draws < 4 * N + 16) {
                probe();
                ++draws;
            }
        }
        particle.weight_set().read_weight(&w[0]);
        double mx = 0, my = 0;
        for (size_t i = 0; i < N; ++i) {
            mx = mx + w[i] * v.state(i, 0);
            my = my + w[i] * v.state(i, 1);
        }
        if (rec) {
            rec[t].ess = ess;
            rec[t].log_z = log_z;
            rec[t].mean_x = mx;
            rec[t].mean_y = my;
            rec[t].resampled = resampled ? 1 : 0;
            rec[t].draws = draws;
        }
    }
    if (state_out)
        for (size_t i = 0; i < N; ++i)
            for (int j = 0; j < 4; ++j) state_out[4 * i + j] = particle.value().state(i, j);
    if (lw_out) particle.weight_set().read_log_weight(lw_out);
    return 0;
}

// The twelve integer-to-uniform conversions of include/vsmc/rng/u01.h:
// out holds 8 values per 32-bit input (closed_closed, closed_open,
// open_closed, open_open; the _24 four as float then the _53 four) and 4 per
// 64-bit input (the _64_53 four in the same order).
void ref_u01_32(const uint32_t *u, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) {
        double *o = out + 8 * i;
        o[0] = u01_closed_closed_32_24(u[i]);
        o[1] = u01_closed_open_32_24(u[i]);
        o[2] = u01_open_closed_32_24(u[i]);
        o[3] = u01_open_open_32_24(u[i]);
        o[4] = u01_closed_closed_32_53(u[i]);
        o[5] = u01_closed_open_32_53(u[i]);
        o[6] = u01_open_closed_32_53(u[i]);
        o[7] = u01_open_open_32_53(u[i]);
    }
}
void ref_u01_64(const uint64_t *u, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) {
        double *o = out + 4 * i;
        o[0] = u01_closed_closed_64_53(u[i]);
        o[1] = u01_closed_open_64_53(u[i]);
        o[2] = u01_open_closed_64_53(u[i]);
        o[3] = u01_open_open_64_53(u[i]);
    }
}

}  // extern "C"
