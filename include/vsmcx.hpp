// vsmcx.hpp -- C++ host layer over the C ABI in vsmcx.h, shaped like vSMC's
// own host API so that vSMC user code can hand the hot path to the GPU.
//
//   vsmcx::Context                RAII over vsmcx_ctx (one HIP stream, HBM budget)
//   vsmcx::ResampleScheme         vSMC's enum (include/vsmc/resample/common.hpp:124-141)
//   vsmcx::HipResample            a functor with the signature of
//                                 Particle<T>::resample_type
//                                 (include/vsmc/core/particle.hpp:61-62); assign it
//                                 with Sampler::resample_scheme(resample_type)
//                                 (include/vsmc/core/sampler.hpp:235-264)
//   vsmcx::Threefry4x64State      engine state transfer in the text layout of
//                                 ThreefryEngine's stream operators
//                                 (include/vsmc/rng/threefry.hpp:424-463)
//   vsmcx::Philox4x32State        the same for PhiloxEngine, whose operator>>
//                                 reads its fields in another order than its
//                                 operator<< writes them
//                                 (include/vsmc/rng/philox.hpp:455-484)
//   vsmcx::Filter                 a device-resident bootstrap filter of the pf
//                                 model: a whole Sampler run (init, iterate,
//                                 histories) in HBM
//   vsmcx::StateHIP<Dim, FP>      a value type T for vsmc::Particle<T> whose
//                                 particles live in HBM; its copy(N, copy_from)
//                                 is StateMatrix::copy
//                                 (include/vsmc/core/state_matrix.hpp:285-300)
//                                 as a device gather
//   vsmcx::PfSampler              a Sampler<T>-shaped driver of the device
//                                 filter: initialize(data file), iterate(n),
//                                 Sampler's histories and its print table
//                                 (include/vsmc/core/sampler.hpp:593-721), as
//                                 the pf example saves it and example/pf/pf.R
//                                 reads it
//   vsmcx::CvMoveHIP              the pf example's init and move
//                                 (example/pf/include/pf_smp.hpp:150-238) over a
//                                 StateHIP<4, FP>, assignable to Sampler's
//                                 init_type and move_type
//                                 (include/vsmc/core/sampler.hpp:65-69); it
//                                 updates vSMC's own WeightSet on the host
//
// Header only; link with libvsmcx.so.  Errors throw vsmcx::Error, which
// carries the C ABI's status code and message (vSMC's runtime asserts throw
// as well when VSMC_RUNTIME_ASSERT_AS_EXCEPTION is set,
// include/vsmc/internal/assert.hpp:57-75).
#ifndef VSMCX_HPP
#define VSMCX_HPP

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vsmcx.h"

namespace vsmcx {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc)
{
    if (rc != VSMCX_OK) throw Error(rc, vsmcx_last_error());
}

class Context {
  public:
    explicit Context(int device = 0, std::uint64_t hbm_budget = 0)
    {
        check(vsmcx_ctx_create(device, hbm_budget, &h_));
    }
    ~Context()
    {
        if (h_) vsmcx_ctx_destroy(h_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    vsmcx_ctx *get() const { return h_; }
    void sync() { check(vsmcx_ctx_sync(h_)); }
    std::uint64_t allocated() const
    {
        std::uint64_t b = 0;
        check(vsmcx_ctx_allocated(h_, &b));
        return b;
    }

  private:
    vsmcx_ctx *h_ = nullptr;
};

enum ResampleScheme {
    Multinomial = VSMCX_MULTINOMIAL,
    Residual = VSMCX_RESIDUAL,
    Stratified = VSMCX_STRATIFIED,
    Systematic = VSMCX_SYSTEMATIC,
    ResidualStratified = VSMCX_RESIDUAL_STRATIFIED,
    ResidualSystematic = VSMCX_RESIDUAL_SYSTEMATIC
};

// Threefry4x64 state: counter, key, the 4-word buffer and the index into it.
// vSMC's engine prints "ctr[4] par[5] buffer[4] index" (par = key words and
// their parity word), which is how the state crosses to and from a
// vsmc::Threefry4x64 without touching its private members.
struct Threefry4x64State : vsmcx_threefry4x64_state {
    std::uint64_t parity = 0;  // par[4], carried through unchanged

    template <typename Engine>
    static Threefry4x64State from(const Engine &eng)
    {
        std::stringstream ss;
        ss << eng;
        Threefry4x64State s;
        for (auto &w : s.ctr) ss >> w;
        for (auto &w : s.key) ss >> w;
        ss >> s.parity;
        for (auto &w : s.buffer) ss >> w;
        ss >> s.index;
        if (!ss) throw Error(VSMCX_ERR_ARG, "cannot read the Threefry4x64 engine state");
        return s;
    }

    template <typename Engine>
    void to(Engine &eng) const
    {
        std::stringstream ss;
        for (auto w : ctr) ss << w << ' ';
        for (auto w : key) ss << w << ' ';
        ss << parity << ' ';
        for (auto w : buffer) ss << w << ' ';
        // trailing blank: the engine's operator>> only commits the state when
        // the stream is still good() after the last field
        ss << index << ' ';
        ss >> eng;
        if (!ss) throw Error(VSMCX_ERR_ARG, "cannot write the Threefry4x64 engine state");
    }
};

// Philox4x32 engine state as text.  vSMC's PhiloxEngine writes
// "ctr key buffer index" (include/vsmc/rng/philox.hpp:455-467) but its
// operator>> reads "buffer ctr key index" (469-484), so a state written by
// one does not come back through the other.  This type reads the written
// layout (from), writes it back in the same layout (write / read round trip),
// and restores an engine through its own operator>> in the order that
// operator reads (to), so the engine continues its stream exactly.
struct Philox4x32State {
    std::uint32_t ctr[4] = {0, 0, 0, 0};
    std::uint32_t key[2] = {0, 0};
    std::uint32_t buffer[4] = {0, 0, 0, 0};
    std::uint32_t index = 4;

    template <typename Engine>
    static Philox4x32State from(const Engine &eng)
    {
        std::stringstream ss;
        ss << eng;
        Philox4x32State s;
        s.read(ss);
        if (!ss) throw Error(VSMCX_ERR_ARG, "cannot read the Philox4x32 engine state");
        return s;
    }

    template <typename Engine>
    void to(Engine &eng) const
    {
        std::stringstream ss;
        for (auto w : buffer) ss << w << ' ';
        for (auto w : ctr) ss << w << ' ';
        for (auto w : key) ss << w << ' ';
        // trailing blank: operator>> commits the state only when the stream
        // is still good() after the last field
        ss << index << ' ';
        ss >> eng;
        if (!ss) throw Error(VSMCX_ERR_ARG, "cannot write the Philox4x32 engine state");
    }

    // the engine's written layout: ctr key buffer index
    template <typename CharT, typename Traits>
    void write(std::basic_ostream<CharT, Traits> &os) const
    {
        for (auto w : ctr) os << w << ' ';
        for (auto w : key) os << w << ' ';
        for (auto w : buffer) os << w << ' ';
        os << index;
    }
    template <typename CharT, typename Traits>
    void read(std::basic_istream<CharT, Traits> &is)
    {
        Philox4x32State s;
        for (auto &w : s.ctr) is >> w;
        for (auto &w : s.key) is >> w;
        for (auto &w : s.buffer) is >> w;
        is >> s.index;
        if (is) *this = s;
    }
};

// Drop-in for Resample<Scheme>::operator()(M, N, rng, weight, replication):
// the replication counts and the advanced engine are the ones vSMC's own
// resampler produces for the same weights and engine state.
class HipResample {
  public:
    HipResample(Context &ctx, ResampleScheme scheme) : ctx_(&ctx), scheme_(scheme) {}

    template <typename Rng, typename IntType>
    void operator()(std::size_t M, std::size_t N, Rng &rng, const double *weight,
                    IntType *replication) const
    {
        Threefry4x64State s = Threefry4x64State::from(rng);
        rep_.resize(M);
        check(vsmcx_resample(ctx_->get(), (int)scheme_, M, N, &s, weight, rep_.data(), nullptr));
        for (std::size_t i = 0; i < M; ++i) replication[i] = static_cast<IntType>(rep_[i]);
        s.to(rng);
    }

    // Replication and, in the same call, the cfrp_trans layout
    // (include/vsmc/resample/common.hpp:87-118) for M == N.
    template <typename Rng, typename IntType>
    void operator()(std::size_t N, Rng &rng, const double *weight, IntType *replication,
                    IntType *copy_from) const
    {
        Threefry4x64State s = Threefry4x64State::from(rng);
        rep_.resize(N);
        cf_.resize(N);
        check(vsmcx_resample(ctx_->get(), (int)scheme_, N, N, &s, weight, rep_.data(),
                             cf_.data()));
        for (std::size_t i = 0; i < N; ++i) {
            replication[i] = static_cast<IntType>(rep_[i]);
            copy_from[i] = static_cast<IntType>(cf_[i]);
        }
        s.to(rng);
    }

  private:
    Context *ctx_;
    ResampleScheme scheme_;
    mutable std::vector<std::uint64_t> rep_, cf_;
};

// The context StateHIP values use when vsmc::Particle<T> constructs them with
// T(N) alone: device 0, created on first use, kept for the life of the process.
inline Context &default_context()
{
    static Context *c = new Context(0);
    return *c;
}

// A vSMC value type with its particles in HBM: N rows of Dim values, FP32 or
// FP64 (FP).  vsmc::Particle<StateHIP<Dim, FP>> calls copy(N, copy_from) after
// every resample (include/vsmc/core/particle.hpp:230-250).
template <std::size_t Dim, typename FP = double>
class StateHIP {
  public:
    typedef std::size_t size_type;
    typedef FP state_type;

    explicit StateHIP(size_type N) : StateHIP(N, default_context()) {}
    StateHIP(size_type N, Context &ctx)
    {
        check(vsmcx_state_create(ctx.get(), N, (std::uint32_t)Dim,
                                 sizeof(FP) == 8 ? VSMCX_FP64_FUSED : VSMCX_MIXED, &h_));
    }
    StateHIP(const StateHIP &o) { check(vsmcx_state_clone(o.h_, &h_)); }
    StateHIP &operator=(const StateHIP &o)
    {
        if (this != &o) {
            vsmcx_state *c = nullptr;
            check(vsmcx_state_clone(o.h_, &c));
            if (h_) vsmcx_state_destroy(h_);
            h_ = c;
        }
        return *this;
    }
    StateHIP(StateHIP &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    StateHIP &operator=(StateHIP &&o) noexcept
    {
        std::swap(h_, o.h_);
        return *this;
    }
    ~StateHIP()
    {
        if (h_) vsmcx_state_destroy(h_);
    }

    size_type size() const
    {
        std::uint64_t n = 0;
        check(vsmcx_state_size(h_, &n, nullptr));
        return (size_type)n;
    }
    static constexpr std::size_t dim() { return Dim; }

    // x[to] = x[copy_from[to]] for every particle
    template <typename IntType>
    void copy(size_type N, const IntType *copy_from)
    {
        std::vector<std::uint64_t> cf(copy_from, copy_from + N);
        check(vsmcx_state_copy(h_, N, cf.data()));
    }

    // host access, synchronous: all rows (N x Dim, row-major)
    std::vector<double> rows() const
    {
        std::vector<double> r(size() * Dim);
        check(vsmcx_state_read(h_, r.data()));
        return r;
    }
    void rows(const std::vector<double> &r) { check(vsmcx_state_write(h_, r.data())); }
    // column pos of every particle, as StateMatrix::read_state
    template <typename OutputIter>
    void read_state(std::size_t pos, OutputIter first) const
    {
        const std::vector<double> r = rows();
        for (size_type i = 0; i != r.size() / Dim; ++i, ++first) *first = r[i * Dim + pos];
    }
    // the current rows on the device, for user kernels
    void *data() const
    {
        void *p = nullptr;
        check(vsmcx_state_data(h_, &p));
        return p;
    }
    vsmcx_state *get() const { return h_; }

  private:
    vsmcx_state *h_ = nullptr;
};

// Model and run description of the device-resident pf filter; the defaults are
// the pf example's (example/pf/include/pf_smp.hpp:96-112,150-238).
struct FilterConfig {
    std::uint64_t n = 10000;
    std::uint64_t horizon = 100;
    int precision = VSMCX_MIXED;
    ResampleScheme scheme = Systematic;
    double threshold = 0.5;  // Sampler::resample_threshold (core/sampler.hpp:270-271)
    std::uint64_t seed = 42, resample_seed = 4242;
    double dt = 0.1, sd_pos = 0.1, sd_vel = 0.5, prior_sd_pos = 2.0, prior_sd_vel = 1.0;
    double obs_scale = 0.1, nu = 10.0;
};

class Filter {
  public:
    Filter(Context &ctx, const FilterConfig &c)
    {
        vsmcx_filter_desc d{};
        d.n = c.n;
        d.horizon = c.horizon;
        d.precision = c.precision;
        d.scheme = (int)c.scheme;
        d.threshold = c.threshold;
        d.seed = c.seed;
        d.resample_seed = c.resample_seed;
        d.dt = c.dt;
        d.sd_pos = c.sd_pos;
        d.sd_vel = c.sd_vel;
        d.prior_sd_pos = c.prior_sd_pos;
        d.prior_sd_vel = c.prior_sd_vel;
        d.obs_scale = c.obs_scale;
        d.nu = c.nu;
        d.llh_const = llh_const(c.nu, c.obs_scale);
        check(vsmcx_filter_create(ctx.get(), &d, &h_));
    }
    // a filter resumed from a checkpoint that save() wrote (vsmcx_filter_load)
    Filter(Context &ctx, const char *checkpoint)
    {
        check(vsmcx_filter_load(ctx.get(), checkpoint, &h_));
    }
    ~Filter()
    {
        if (h_) vsmcx_filter_destroy(h_);
    }
    void save(const char *path) const { check(vsmcx_filter_save(h_, path)); }
    Filter(const Filter &) = delete;
    Filter &operator=(const Filter &) = delete;

    // observations: T rows of (x, y)
    void observations(const std::vector<double> &xy)
    {
        check(vsmcx_filter_set_obs(h_, xy.size() / 2, xy.data()));
    }
    // Sampler::initialize + iterate(n - 1) on the first call, iterate(n) after
    void iterate(std::uint64_t n = 1) { check(vsmcx_filter_run(h_, n)); }
    void finish() { check(vsmcx_filter_finish(h_)); }
    std::uint64_t iter_size() const
    {
        std::uint64_t t = 0;
        check(vsmcx_filter_steps(h_, &t));
        return t;
    }
    std::vector<vsmcx_step_stats> history() const
    {
        std::vector<vsmcx_step_stats> h(iter_size());
        check(vsmcx_filter_history(h_, h.size(), h.data()));
        return h;
    }
    vsmcx_filter *get() const { return h_; }

    // Student-t log normaliser per coordinate: lgamma((nu+1)/2) - lgamma(nu/2)
    // - log(nu * pi) / 2 - log(scale)
    static double llh_const(double nu, double scale);

  private:
    vsmcx_filter *h_ = nullptr;
};

// A Sampler<T>-shaped driver of the device-resident pf filter
// (include/vsmc/core/sampler.hpp).  initialize() is Sampler::initialize(param)
// with the pf example's param, the observation file of (x, y) pairs
// (example/pf/include/pf_smp.hpp:113-125, 175-176): the init step, its
// resample and the monitor (do_init, core/sampler.hpp:773-781).  iterate(n)
// is n further steps (do_iter, 783-792).  The histories have Sampler's
// meanings (do_resample, 816-822): size and ESS before the resample, whether
// it resampled, the accept counts of the init and of the move (the pf init and
// move return 1 per particle, so N), and the "pos" monitor (pos.x, pos.y)
// after the resample (MonitorMCMC, the pf example's cv_est).  print() and
// operator<< write Sampler::print's table (core/sampler.hpp:679-721):
//   Size Resampled Accept.0 ESS pos.x pos.y
// every field followed by the separator, each row ended by '\n', numbers in
// the stream's own format -- the file the pf example writes
// (example/pf/include/pf_mpi_do.hpp:58-64) and example/pf/pf.R:53-111 reads.
class PfSampler {
  public:
    // Sampler<T>(N, scheme, threshold) (core/sampler.hpp:101-104); the model
    // constants, seeds and precision come from base (its n, scheme and
    // threshold are replaced)
    PfSampler(Context &ctx, std::size_t N, ResampleScheme scheme, double threshold = 0.5,
              const FilterConfig &base = FilterConfig())
        : ctx_(&ctx), cfg_(base)
    {
        cfg_.n = N;
        cfg_.scheme = scheme;
        cfg_.threshold = threshold;
    }

    // Sampler::initialize with the pf data file: one (x, y) pair per line,
    // read to the end of the file; the filter's horizon is the number of pairs
    PfSampler &initialize(const char *file)
    {
        std::ifstream in(file);
        if (!in) throw Error(VSMCX_ERR_ARG, std::string("cannot open ") + file);
        std::vector<double> xy;
        double x, y;
        while (in >> x >> y) {
            xy.push_back(x);
            xy.push_back(y);
        }
        return initialize(xy);
    }
    // the same from T rows of (x, y)
    PfSampler &initialize(const std::vector<double> &xy)
    {
        if (xy.size() < 2) throw Error(VSMCX_ERR_ARG, "no observations");
        cfg_.horizon = xy.size() / 2;
        filter_.reset();  // Sampler::do_reset: a new run
        filter_.reset(new Filter(*ctx_, cfg_));
        filter_->observations(xy);
        filter_->iterate(1);
        refresh();
        return *this;
    }
    PfSampler &iterate(std::size_t num = 1)
    {
        if (!filter_) throw Error(VSMCX_ERR_STATE, "iterate before initialize");
        if (num) filter_->iterate(num);
        refresh();
        return *this;
    }

    std::size_t size() const { return (std::size_t)cfg_.n; }
    std::size_t iter_size() const { return hist_.size(); }
    std::size_t size_history(std::size_t iter) const
    {
        (void)at(iter);
        return size();
    }
    double ess_history(std::size_t iter) const { return at(iter).ess; }
    bool resampled_history(std::size_t iter) const { return at(iter).resampled != 0; }
    std::size_t accept_history(std::size_t id, std::size_t iter) const
    {
        if (id != 0) throw Error(VSMCX_ERR_ARG, "the pf sampler has one move");
        (void)at(iter);
        return size();
    }
    // the "pos" monitor, i = 0 (pos.x) or 1 (pos.y)
    double monitor_record(std::size_t i, std::size_t iter) const
    {
        return i == 0 ? at(iter).mean_x : at(iter).mean_y;
    }
    // running log-evidence after iter (not part of Sampler's table)
    double log_z_history(std::size_t iter) const { return at(iter).log_z; }
    Filter &filter() { return *filter_; }

    template <typename CharT, typename Traits>
    std::basic_ostream<CharT, Traits> &print(std::basic_ostream<CharT, Traits> &os,
                                             char sepchar = '\t') const
    {
        if (iter_size() == 0 || !os.good()) return os;
        os << "Size" << sepchar << "Resampled" << sepchar << "Accept.0" << sepchar;
        os << "ESS" << sepchar << "pos.x" << sepchar << "pos.y" << sepchar << '\n';
        for (std::size_t iter = 0; iter != iter_size(); ++iter) {
            os << size_history(iter) << sepchar << (std::size_t)resampled_history(iter) << sepchar
               << accept_history(0, iter) << sepchar;
            os << ess_history(iter) << sepchar << monitor_record(0, iter) << sepchar
               << monitor_record(1, iter) << sepchar;
            os << '\n';
        }
        return os;
    }

  private:
    Context *ctx_;
    FilterConfig cfg_;
    std::unique_ptr<Filter> filter_;
    std::vector<vsmcx_step_stats> hist_;

    // the last step's monitor is evaluated by the next step's move; finish()
    // evaluates it now (the next move recomputes it)
    void refresh()
    {
        filter_->finish();
        hist_ = filter_->history();
    }
    const vsmcx_step_stats &at(std::size_t iter) const
    {
        if (iter >= hist_.size()) throw Error(VSMCX_ERR_ARG, "iteration out of range");
        return hist_[iter];
    }
};

// Sampler's operator<< (core/sampler.hpp:950-953)
template <typename CharT, typename Traits>
inline std::basic_ostream<CharT, Traits> &operator<<(std::basic_ostream<CharT, Traits> &os,
                                                      const PfSampler &sampler)
{
    return sampler.print(os);
}

// The pf example's init and move over a StateHIP<4, FP>: the prior draw (as
// init_type, size_t(Particle<T> &, void *)) or the transition at iteration
// iter (as move_type, size_t(size_t, Particle<T> &)) of every particle on the
// device, then the log-likelihood increments into the Particle's own
// WeightSet: set_log_weight at init, add_log_weight after
// (include/vsmc/core/weight_set.hpp:297-331).  Observation iter is row iter of
// the (x, y) rows given.
class CvMoveHIP {
  public:
    CvMoveHIP(const FilterConfig &c, const std::vector<double> &obs_xy) : obs_(obs_xy)
    {
        d_.n = c.n;
        d_.horizon = c.horizon;
        d_.precision = c.precision;
        d_.scheme = (int)c.scheme;
        d_.threshold = c.threshold;
        d_.seed = c.seed;
        d_.resample_seed = c.resample_seed;
        d_.dt = c.dt;
        d_.sd_pos = c.sd_pos;
        d_.sd_vel = c.sd_vel;
        d_.prior_sd_pos = c.prior_sd_pos;
        d_.prior_sd_vel = c.prior_sd_vel;
        d_.obs_scale = c.obs_scale;
        d_.nu = c.nu;
        d_.llh_const = Filter::llh_const(c.nu, c.obs_scale);
    }

    template <typename P>
    std::size_t operator()(P &particle, void *)
    {
        step(particle, 0);
        particle.weight_set().set_log_weight(inc_.data());
        return 0;
    }
    template <typename P>
    std::size_t operator()(std::size_t iter, P &particle)
    {
        step(particle, iter);
        particle.weight_set().add_log_weight(inc_.data());
        return 0;
    }
    // the increments of the last call
    const std::vector<double> &increments() const { return inc_; }

  private:
    template <typename P>
    void step(P &particle, std::size_t iter)
    {
        if (2 * iter + 1 >= obs_.size()) throw Error(VSMCX_ERR_ARG, "no observation for this iteration");
        inc_.resize(particle.size());
        check(vsmcx_state_cv_move(particle.value().get(), &d_, iter, &obs_[2 * iter], inc_.data()));
    }

    vsmcx_filter_desc d_{};
    std::vector<double> obs_;
    std::vector<double> inc_;
};
typedef CvMoveHIP CvInitHIP;

}  // namespace vsmcx

#include <cmath>

inline double vsmcx::Filter::llh_const(double nu, double scale)
{
    return std::lgamma((nu + 1) / 2) - std::lgamma(nu / 2) -
           0.5 * std::log(nu * 3.14159265358979323846) - std::log(scale);
}

#endif  // VSMCX_HPP
