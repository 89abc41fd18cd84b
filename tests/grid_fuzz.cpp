// Host-side fuzz of the frequency-grid entry points (brpa_dai_histogram,
// brpa_freq_grid_ex, brpa_optimize_freq_grid, brpa_dai_max_thresh,
// brpa_opt_quad_scaling, brpa_optimize_freq_quad) on random and degenerate
// orbital energies: equal occupied energies (zero-width gap range), histogram
// caps below the smallest gap, non-positive gaps, one-orbital sets.  Built by
// tests/test_grid_sanitized.py with AddressSanitizer and UndefinedBehavior-
// Sanitizer (float-cast-overflow included), which stop the run at the first
// finding.  Prints "ok <n> err <n>": calls that succeeded and calls that
// returned an error code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "brpa.h"

int main(int argc, char **argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 40;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(0, 1);
    int n_ok = 0, n_err = 0;
    auto tally = [&](int rc) { rc == BRPA_OK ? ++n_ok : ++n_err; };
    double vals[100], w[100], fx[64], fw[64], info[3], ofx[24], ofw[24];
    for (int it = 0; it < iters; ++it) {
        const int no = 1 + (int)(g() % 6), nv = 1 + (int)(g() % 8);
        std::vector<double> eo(no), ev(nv);
        for (auto &x : eo) x = -20 * U(g) * (it % 3 == 0 ? 1.0 : 0.1);   // some core orbitals
        for (auto &x : ev) x = (it % 7 == 0) ? 0.3 : 3 * U(g) - (it % 11 == 0 ? 2.0 : 0.0);
        std::sort(eo.begin(), eo.end());
        std::sort(ev.begin(), ev.end());
        if (it % 5 == 0)
            for (auto &x : eo) x = eo.back();                     // every active gap equal
        const double cap = (it % 4 == 0) ? 0.2 + U(g) : 0.0;      // may sit below the smallest gap
        tally(brpa_dai_histogram(eo.data(), no, ev.data(), nv, -3.0, cap, vals, w));
        const int nf = 2 + (int)(g() % 20);
        tally(brpa_freq_grid_ex(eo.data(), no, ev.data(), nv, nf, -3.0, cap, fx, fw, info));
        int64_t n = 0;
        tally(brpa_optimize_freq_grid(eo.data(), no, ev.data(), nv, 1e-6, 1e-4, 2, 24, ofx, ofw,
                                      &n, info));
        int64_t nos[2] = {no, no}, nvs[2] = {nv, nv};
        std::vector<double> eo2(eo), ev2(ev);
        eo2.insert(eo2.end(), eo.begin(), eo.end());
        ev2.insert(ev2.end(), ev.begin(), ev.end());
        double t = 0.0;
        tally(brpa_dai_max_thresh(2, eo2.data(), nos, ev2.data(), nvs, -3.0, &t));
        if (brpa_dai_histogram(eo.data(), no, ev.data(), nv, -3.0, 0.0, vals, w) == BRPA_OK) {
            double v2[200], w2[200];
            for (int k = 0; k < 100; ++k) {
                v2[k] = vals[k];
                w2[k] = w[k];
                v2[100 + k] = 2.0 * vals[k];
                w2[100 + k] = w[k];
            }
            tally(brpa_opt_quad_scaling(nf, 2, v2, w2, fx, fw, info));
            tally(brpa_optimize_freq_quad(2, v2, w2, 1e-6, 1e-4, 2, 24, ofx, ofw, &n, info));
        }
    }
    std::printf("ok %d err %d\n", n_ok, n_err);
    return 0;
}
