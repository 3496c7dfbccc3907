// Host-side check of vsmc_amd/csrc/fold_core.hpp: the tiled transducer
// evaluation (classify with an approximate prefix, compose regular segments,
// chain specials with real FP64 adds) must equal the sequential left fold
// bit for bit.  Built and run by tests/test_fold_core.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../vsmc_amd/csrc/fold_core.hpp"

using namespace vsmcx;

static int check_case(const std::vector<double> &x, int tile, bool perturb_approx, int *nspec_out)
{
    const size_t n = x.size();
    // sequential left fold (the definition)
    std::vector<double> ref(n);
    double a = 0;
    for (size_t i = 0; i < n; ++i) {
        a += x[i];
        ref[i] = a;
    }
    double eps_rel, eps_abs;
    fold_eps(n, &eps_rel, &eps_abs);
    // approximate prefix: tile sums in pairwise order, then within-tile running sums
    const size_t ntiles = (n + tile - 1) / tile;
    std::vector<double> tsum(ntiles, 0.0), tstart(ntiles, 0.0);
    for (size_t t = 0; t < ntiles; ++t) {
        std::vector<double> v(x.begin() + t * tile, x.begin() + std::min(n, (t + 1) * tile));
        while (v.size() > 1) {
            std::vector<double> w;
            for (size_t k = 0; k + 1 < v.size(); k += 2) w.push_back(v[k] + v[k + 1]);
            if (v.size() % 2) w.push_back(v.back());
            v = w;
        }
        tsum[t] = v.empty() ? 0 : v[0];
        if (perturb_approx) tsum[t] *= 1 + 1e-13;
    }
    for (size_t t = 1; t < ntiles; ++t) tstart[t] = tstart[t - 1] + tsum[t - 1];
    // classification + per-element transducers
    std::vector<int> special(n), bin(n);
    std::vector<double> S(n);
    for (size_t t = 0; t < ntiles; ++t) {
        double s = tstart[t];
        for (size_t i = t * tile; i < std::min(n, (t + 1) * tile); ++i) {
            const double sp = s;
            s += x[i];
            S[i] = s;
            const int b0 = certain_binade(sp, eps_rel, eps_abs);
            const int b1 = certain_binade(s, eps_rel, eps_abs);
            special[i] = (i == 0) || b0 != b1 || b1 == kNoBinade;
            bin[i] = b1;
        }
    }
    // walk: compose regular runs, add specials exactly; verify every value
    int nspec = 0;
    double v = 0;
    Tr run = tr_identity();
    double run_start = 0;
    for (size_t i = 0; i < n; ++i) {
        double got;
        if (special[i]) {
            ++nspec;
            double vb;
            if (!tr_apply(run, run_start, &vb)) return 1;
            v = vb + x[i];
            run = tr_identity();
            run_start = v;
            got = v;
        } else {
            run = tr_compose(run, tr_element(x[i], bin[i]));
            if (!tr_apply(run, run_start, &got)) return 2;
        }
        if (got != ref[i]) {
            std::fprintf(stderr, "mismatch at %zu: %.17g vs %.17g (special %d)\n", i, got, ref[i],
                         special[i]);
            return 3;
        }
    }
    *nspec_out = nspec;
    return 0;
}

int main()
{
    std::mt19937_64 g(12345);
    int failures = 0, cases = 0;
    long total_spec = 0;
    for (int rep = 0; rep < 60; ++rep) {
        const size_t n = 1 + g() % 200000;
        std::vector<double> x(n);
        const int kind = rep % 6;
        std::uniform_real_distribution<double> U(0, 1);
        std::normal_distribution<double> Z(0, 1);
        for (size_t i = 0; i < n; ++i) {
            switch (kind) {
            case 0: x[i] = U(g); break;                             // flat
            case 1: x[i] = std::exp(3 * Z(g)); break;               // heavy tail
            case 2: x[i] = (g() % 3 == 0) ? 0.0 : U(g); break;      // zeros
            case 3: x[i] = std::ldexp(1.0, -(int)(g() % 60)); break; // powers of two: ties
            case 4: x[i] = std::exp(-0.01 * (double)i); break;      // decaying
            default: x[i] = std::exp(20 * Z(g)); break;             // huge dynamic range
            }
        }
        if (kind != 3) {
            double s = 0;
            for (double v : x) s += v;
            for (double &v : x) v /= s;  // normalised weights
        }
        for (int tile : {256, 4096}) {
            int ns = 0;
            int rc = check_case(x, tile, rep % 2 == 1, &ns);
            ++cases;
            total_spec += ns;
            if (rc) {
                std::fprintf(stderr, "case %d kind %d n %zu tile %d failed rc=%d\n", rep, kind, n,
                             tile, rc);
                ++failures;
            }
        }
    }
    // exhaustive-ish tiny dyadic cases with many ties
    for (int rep = 0; rep < 20000; ++rep) {
        const size_t n = 1 + g() % 9;
        std::vector<double> x(n);
        for (size_t i = 0; i < n; ++i) x[i] = std::ldexp((double)(g() % 9), -(int)(g() % 56));
        int ns = 0;
        if (check_case(x, 4, false, &ns)) ++failures;
        ++cases;
    }
    std::printf("cases %d failures %d specials %ld\n", cases, failures, total_spec);
    return failures ? 1 : 0;
}
