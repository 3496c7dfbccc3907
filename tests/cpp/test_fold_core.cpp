// Host-side check of vsmc_amd/csrc/fold_core.hpp: the segmented evaluation
// the device runs (classify every element against an approximate prefix,
// per-thread runs with the single-binade fast path, segment summaries combined
// in tile/thread groupings, one real FP64 add per special element, then every
// value from its run) must equal the sequential left fold
// a_i = fl(a_{i-1} + x_i) bit for bit.
// Built and run by tests/test_fold_core.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../vsmc_amd/csrc/fold_core.hpp"

using namespace vsmcx;

// Compose a range of Seg values in a random binary grouping (associativity is
// what lets tiles and threads combine in any order).
static Seg fold_random(const std::vector<Seg> &v, size_t lo, size_t hi, std::mt19937_64 &g)
{
    if (hi - lo == 0) return seg_identity();
    if (hi - lo == 1) return v[lo];
    const size_t mid = lo + 1 + g() % (hi - lo - 1);
    return seg_op(fold_random(v, lo, mid, g), fold_random(v, mid, hi, g));
}

static bool same(const Seg &a, const Seg &b)
{
    return a.cnt == b.cnt && a.tr.d0 == b.tr.d0 && a.tr.d1 == b.tr.d1 && a.tr.pp == b.tr.pp &&
           a.tr.b == b.tr.b;
}

static int check_case(const std::vector<double> &x, int items, int tile, double perturb,
                      std::mt19937_64 &g, long *nspec_out)
{
    const size_t n = x.size();
    std::vector<double> ref(n);
    double a = 0;
    for (size_t i = 0; i < n; ++i) {
        a += x[i];
        ref[i] = a;
    }
    double er, ea;
    fold_eps(n, &er, &ea);
#ifdef VSMCX_TEST_ZERO_EPS
    er = 0;  // a broken bound, to show that it is caught
    ea = 0;
#endif
    // The passes of kernels_fold.hip on the host: tiles of `tile` elements,
    // threads of `items` consecutive elements (tail padded with zeros).
    const size_t ntiles = (n + tile - 1) / tile;
    const int nthr = tile / items;
    auto X = [&](size_t i) { return i < n ? x[i] : 0.0; };
    // pass 1-2: tile sums and approximate tile starts (perturbed within budget)
    std::vector<double> ts(ntiles);
    double start = 0;
    for (size_t t = 0; t < ntiles; ++t) {
        ts[t] = start;
        double tsum = 0;
        for (int e = 0; e < tile; ++e) tsum += X(t * tile + e);
        start = (start + tsum) * (1 + perturb);
    }
    // pass 3: per-thread runs, tile summaries, tile-local special records
    std::vector<double> s_thr(ntiles * nthr);
    std::vector<int> fast(ntiles * nthr);
    std::vector<Seg> gthr(ntiles * nthr), agg(ntiles);
    std::vector<SpecRec> spec(ntiles * tile);
    for (size_t t = 0; t < ntiles; ++t) {
        double run = 0;
        Seg tl = seg_identity();
        for (int j = 0; j < nthr; ++j) {
            const size_t i0 = t * tile + (size_t)j * items;
            const size_t id = t * nthr + j;
            s_thr[id] = ts[t] + run;
            double p = 0;
            for (int k = 0; k < items; ++k) p += X(i0 + k);
            run += p;
            auto xat = [&](int k) { return X(i0 + k); };
            gthr[id] = run_seg(xat, items, s_thr[id], er, ea, &fast[id]);
            if (gthr[id].cnt) {
                Seg pre = tl;
                double s = s_thr[id];
                for (int k = 0; k < items; ++k) {
                    const double sb = s;
                    s = sb + X(i0 + k);
                    const Seg el = elem_seg(X(i0 + k), sb, s, er, ea);
                    if (el.cnt) {
                        spec[t * tile + pre.cnt].run = pre.tr;
                        spec[t * tile + pre.cnt].x = X(i0 + k);
                    }
                    pre = seg_op(pre, el);
                }
            }
            tl = seg_op(tl, gthr[id]);
        }
        agg[t] = tl;
        if (!same(fold_random(gthr, t * nthr, (t + 1) * nthr, g), tl)) return 4;
    }
    // pass 4: tile prefixes, specials in fold order, the chain
    std::vector<Seg> pre(ntiles);
    Seg acc = seg_identity();
    for (size_t t = 0; t < ntiles; ++t) {
        pre[t] = acc;
        acc = seg_op(acc, agg[t]);
    }
    if (!same(fold_random(agg, 0, ntiles, g), acc)) return 4;
    const uint64_t nspec = acc.cnt;
    std::vector<double> gval(nspec + 1);
    double v = 0;
    uint64_t gi = 0;
    for (size_t t = 0; t < ntiles; ++t)
        for (uint64_t k = 0; k < agg[t].cnt; ++k, ++gi) {
            SpecRec r = spec[t * tile + k];
            if (k == 0) r.run = tr_compose(pre[t].tr, r.run);
            double vb;
            if (!tr_apply(r.run, v, &vb)) return 1;
            v = vb + r.x;
            gval[gi] = v;
        }
    double tot;
    if (!tr_apply(acc.tr, v, &tot)) return 6;
    // pass 5: every value, as the apply pass computes it: the exact value
    // before each thread's run from its prefix summary, then the run's own
    // fold steps; every run must end where the next run starts (the last run:
    // at the fold total), or the pass reports a broken bound (rc 2)
    double end_prev = 0.0;  // exact value before the first run
    for (size_t t = 0; t < ntiles; ++t) {
        Seg tp = pre[t];
        for (int j = 0; j < nthr; ++j) {
            const size_t i0 = t * tile + (size_t)j * items;
            const size_t id = t * nthr + j;
            const RunStart st = run_start(tp, [&](uint64_t c) { return gval[c]; });
            if (!st.ok || !same_value(st.a_prev, end_prev)) return 2;
            double a = st.a_prev;
            for (int k = 0; k < items; ++k) {
                a = a + X(i0 + k);
                const size_t i = i0 + k;
                if (i < n && !same_value(a, ref[i])) {
                    std::fprintf(stderr, "mismatch at %zu: %.17g vs %.17g\n", i, a, ref[i]);
                    return 3;
                }
            }
            end_prev = a;
            tp = seg_op(tp, gthr[id]);
        }
    }
    if (!same_value(end_prev, tot)) return 2;
    if (tot != ref[n - 1] && !(std::isnan(tot) && std::isnan(ref[n - 1]))) return 3;
    *nspec_out = (long)nspec;
    return 0;
}

// The fast path's units from its own fold steps started at 2^b, against the
// units counted element by element on the grid (q + [fraction > 1/2], ties
// refused), over runs that sit anywhere in their binade, with exact halves,
// zeros, values below half a grid unit and binades down to the subnormals.
static int check_fast_units(std::mt19937_64 &g)
{
    int bad = 0;
    std::uniform_real_distribution<double> U(0, 1);
    for (int rep = 0; rep < 400000; ++rep) {
        const int b = (rep % 7 == 0) ? -1000 - (int)(g() % 23) : (int)(g() % 80) - 60;
        const double lo = std::ldexp(1.0, b);
        // a start inside the binade and eight increments that keep the run in it
        const double s = lo * (1.0 + 0.5 * U(g));
        const double room = std::ldexp(1.0, b + 1) - s;
        double x[8];
        for (int k = 0; k < 8; ++k) {
            const int kind = (int)(g() % 6);
            const double grid = std::ldexp(1.0, b - 52);
            if (kind == 0) x[k] = 0.0;
            else if (kind == 1) x[k] = grid * (0.5 + (double)(g() % 64));  // exact tie
            else if (kind == 2) x[k] = grid * 0.25 * U(g);                 // rounds away
            else x[k] = room / 9 * U(g);
        }
        int fb = kNoBinade;
        Seg r;
        const bool fast = run_seg_fast([&](int k) { return x[k]; }, 8, s, 0.0, 0.0, &fb, &r);
        uint64_t d = 0;
        int tie = 0;
        for (int k = 0; k < 8; ++k) {
            int kind;
            d += grid_units(x[k], binade_of(s), &kind) + (kind == 1);
            tie |= kind == 2;
        }
        double se = s;
        for (int k = 0; k < 8; ++k) se += x[k];
        const bool one_binade = binade_of(s) == binade_of(se) && binade_of(s) < 1024;
        if (fast != (one_binade && !tie)) ++bad;
        if (fast && (fb != binade_of(s) || r.tr.d0 != d || r.tr.d1 != d)) ++bad;
    }
    return bad;
}

int main()
{
    std::mt19937_64 g(12345);
    int failures = 0, cases = 0;
    long total_spec = 0, max_spec = 0;
    std::uniform_real_distribution<double> U(0, 1);
    std::normal_distribution<double> Z(0, 1);
    for (int rep = 0; rep < 90; ++rep) {
        const size_t n = 1 + g() % 200000;
        std::vector<double> x(n);
        const int kind = rep % 9;
        for (size_t i = 0; i < n; ++i) {
            switch (kind) {
            case 0: x[i] = U(g); break;                              // flat
            case 1: x[i] = std::exp(3 * Z(g)); break;                // heavy tail
            case 2: x[i] = (g() % 3 == 0) ? 0.0 : U(g); break;       // zeros
            case 3: x[i] = std::ldexp(1.0, -(int)(g() % 60)); break; // powers of two: ties
            case 4: x[i] = std::exp(-0.01 * (double)i); break;       // decaying
            case 5: x[i] = std::exp(20 * Z(g)); break;               // huge dynamic range
            case 6: x[i] = (i == n / 3) ? 1.0 : 0.0; break;          // one particle
            case 7: x[i] = (i == n / 2) ? 1.0 : 1e-300 * U(g); break;  // one big, rest tiny
            default:  // residual fractions scaled by 1/0: NaN (and an infinity)
                x[i] = (i % 5 == 0) ? 0.0 : (i == n / 2 ? INFINITY : NAN);
                break;
            }
        }
        if (kind != 3 && kind != 8) {
            double s = 0;
            for (double v : x) s += v;
            if (s > 0)
                for (double &v : x) v /= s;  // normalised weights
        }
        for (int tile : {256, 2048}) {
            long ns = 0;
            const int rc = check_case(x, 8, tile, rep % 2 ? 1e-13 : 0.0, g, &ns);
            ++cases;
            total_spec += ns;
            max_spec = std::max(max_spec, ns);
            if (tile == 2048 && rep < 9) std::printf("kind %d n %zu specials %ld\n", kind, n, ns);
            if (rc) {
                std::fprintf(stderr, "case %d kind %d n %zu tile %d failed rc=%d\n", rep, kind, n,
                             tile, rc);
                ++failures;
            }
        }
    }
    // tiny dyadic cases with many ties and exact binade landings
    for (int rep = 0; rep < 20000; ++rep) {
        const size_t n = 1 + g() % 9;
        std::vector<double> x(n);
        for (size_t i = 0; i < n; ++i) x[i] = std::ldexp((double)(g() % 9), -(int)(g() % 56));
        long ns = 0;
        if (check_case(x, 2, 4, 0.0, g, &ns)) ++failures;
        ++cases;
    }
    const int unit_bad = check_fast_units(g);
    if (unit_bad) std::fprintf(stderr, "fast-path units differ in %d runs\n", unit_bad);
    failures += unit_bad;
    std::printf("cases %d failures %d specials total %ld max %ld\n", cases, failures, total_spec,
                max_spec);
    return failures ? 1 : 0;
}
