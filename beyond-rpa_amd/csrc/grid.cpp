// Imaginary-frequency grid of the MO-basis RPA energy (host code).
//
//   ncore            rpa_NCore          rpa/OrbDiffHist.f90:126-146
//   dai_max_thresh   rpa_DaiMaxThresh   rpa/OrbDiffHist.f90:10-39
//   dai_histogram    rpa_DaiHistogram   rpa/OrbDiffHist.f90:42-123
//   gauss_legendre   quad_GaussLegendre rpa/quadratures.f90:418-506 (Newton in theta)
//   casimir_polder   quad_CasimirPolder rpa/quadratures.f90:570-643
//   quad_error       rpa_FreqQuadError  rpa/OptFreqQuad.f90:125-169
//   opt_scaling      rpa_OptQuadScaling rpa/OptFreqQuad.f90:75-122
//   optimize         rpa_OptimizeFreqQuad rpa/OptFreqQuad.f90:12-72
//
// A histogram set is one spin of one system (rpa_driver.f90:118-172 builds one
// per system and spin); the quadrature error of a grid is the worst over the
// sets.  Sets whose occupied orbitals are all frozen carry the reference's
// special values (weights -1), which drop them from every sum.
#include <cmath>
#include <limits>
#include <vector>

#include "internal.h"

namespace {

constexpr int kNBins = BRPA_HISTOGRAM_NBINS;   // RPA_HISTOGRAM_NBINS (rpa_definitions.f90:113)
constexpr double kAlphaMin = 0.1, kAlphaMax = 15.0;
constexpr int kNScan = 15000;          // OptFreqQuad.f90:92-94
constexpr double kPi = 3.14159265358979323846;

// leading occupied orbitals below the threshold (OrbDiffHist.f90:126-146)
int64_t ncore(const double *eocc, int64_t nocc, double thresh) {
    int64_t n = 0;
    while (n < nocc && eocc[n] < thresh) ++n;
    return n;
}

// values/weights: kNBins entries each.  cap: DaiMaxThresh (gaps above it are
// not counted; the histogram's range ends at min(cap, largest active gap)).
int dai_histogram(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                  double core_thresh, double cap, double *values, double *weights) {
    const int64_t nc = ncore(eocc, nocc, core_thresh);
    if (nc >= nocc) {
        // every occupied orbital frozen: values/weights that no sum sees (106-121)
        for (int k = 0; k < kNBins; ++k) {
            values[k] = 0.0;
            weights[k] = -1.0;
        }
        values[0] = std::numeric_limits<double>::max();
        values[kNBins - 1] = 0.0;
        return BRPA_OK;
    }
    const double dmin = evir[0] - eocc[nocc - 1];
    const double dmax = std::fmin(cap, evir[nvir - 1] - eocc[nc]);
    if (!(dmin > 0.0)) return BRPA_ERR_GAP;
    const double width = (dmax - dmin) / kNBins;
    for (int k = 0; k < kNBins; ++k) values[k] = dmin + k * width;
    std::vector<int64_t> count(kNBins, 0);
    int64_t total = 0;
    for (int64_t i = nc; i < nocc; ++i)
        for (int64_t a = 0; a < nvir; ++a) {
            double d = evir[a] - eocc[i];
            if (d <= cap) {
                double t = (d - (dmin + width / 2.0)) / width;
                int64_t k;
                if (!(std::fabs(t) <= (double)kNBins)) {
                    // NaN (all gaps equal: width 0) or beyond the range: the
                    // clamp of the bin rule decides; a NaN quotient goes to the
                    // first bin, where nint of it lands on x86 before max(1, k)
                    k = t > 0.0 ? kNBins : 1;
                } else {
                    // Fortran nint: round half away from zero
                    k = 1 + (int64_t)(t >= 0 ? std::floor(t + 0.5) : -std::floor(-t + 0.5));
                    if (k < 1) k = 1;
                    if (k > kNBins) k = kNBins;
                }
                count[k - 1] += 1;
                total += 1;
            }
        }
    for (int k = 0; k < kNBins; ++k) weights[k] = std::fabs((double)count[k] / (double)total);
    return BRPA_OK;
}

void legendre(double x, int n, double &pn, double &pn1) {
    double p0 = 1.0, p1 = x;
    if (n == 0) {
        pn = 1.0;
        pn1 = 0.0;
        return;
    }
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    pn = p1;
    pn1 = p0;
}

// All n nodes ascending on [-1, 1] and weights.
bool gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    std::vector<double> th(n), nw(n);
    for (int k = 1; k <= n; ++k) th[k - 1] = kPi * (4.0 * k - 1.0) / (4.0 * n + 2.0);
    bool conv = false;
    for (int it = 0; it < 20; ++it) {
        bool all = true;
        for (int k = 0; k < n; ++k) {
            double c = std::cos(th[k]), pn, pn1;
            legendre(c, n, pn, pn1);
            double dp = n * (c * pn - pn1) / std::sin(th[k]);
            nw[k] = th[k] - pn / dp;
            if (!(std::fabs(nw[k] - th[k]) < 1.0e-14)) all = false;
        }
        th = nw;
        if (all) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    x.resize(n);
    w.resize(n);
    // theta ascending -> x descending; store ascending
    for (int k = 0; k < n; ++k) {
        double c = std::cos(th[k]), pn, pn1;
        legendre(c, n, pn, pn1);
        double dp = n * (c * pn - pn1) / std::sin(th[k]);
        x[n - 1 - k] = c;
        w[n - 1 - k] = 2.0 / (dp * dp);
    }
    return true;
}

bool casimir_polder(int n, double alpha, std::vector<double> &x, std::vector<double> &w) {
    std::vector<double> t, wg;
    if (!gauss_legendre(n, t, wg)) return false;
    x.resize(n);
    w.resize(n);
    for (int k = 0; k < n; ++k) {
        x[k] = alpha * (1.0 + t[k]) / (1.0 - t[k]);
        w[k] = 2.0 * wg[k] * alpha / ((1.0 - t[k]) * (1.0 - t[k]));
    }
    return true;
}

// worst quadrature error over nsets histograms (vals/dw: nsets x kNBins)
void quad_error(const double *fx, const double *fw, int n, int64_t nsets, const double *vals,
                const double *dw, double &rel, double &rmsd) {
    double err2max = 0.0;
    rel = 0.0;
    for (int64_t s = 0; s < nsets; ++s) {
        double err2 = 0.0, rels = 0.0;
        for (int l = 0; l < kNBins; ++l) {
            if (dw[s * kNBins + l] > 0.0) {
                const double d = vals[s * kNBins + l];
                const double exact = kPi / (2.0 * d);
                double approx = 0.0;
                for (int k = 0; k < n; ++k) approx += fw[k] / (d * d + fx[k] * fx[k]);
                rels = std::fmax(rels, std::fabs((approx - exact) / exact));
                err2 += dw[s * kNBins + l] * (approx - exact) * (approx - exact);
            }
        }
        err2max = std::fmax(err2max, err2);
        rel = std::fmax(rel, rels);
    }
    rmsd = std::sqrt(err2max);
}

bool opt_scaling(int n, int64_t nsets, const double *vals, const double *dw, double &alpha,
                 double &rel, double &rmsd) {
    std::vector<double> xu, wu;
    if (!casimir_polder(n, 1.0, xu, wu)) return false;
    std::vector<double> fx(n), fw(n);
    double best = std::numeric_limits<double>::max();
    alpha = 0.0;
    rel = 0.0;
    for (int k = 1; k <= kNScan; ++k) {
        double a = kAlphaMin + (kAlphaMax - kAlphaMin) * (double)(k - 1) / (double)(kNScan - 1);
        for (int j = 0; j < n; ++j) {
            fx[j] = a * xu[j];
            fw[j] = a * wu[j];
        }
        double r, m;
        quad_error(fx.data(), fw.data(), n, nsets, vals, dw, r, m);
        if (m < best) {
            best = m;
            rel = r;
            alpha = a;
        }
    }
    rmsd = best;
    return true;
}

bool finite_all(const double *p, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// Grid of nfreqs points on the given histograms (rpa_OptQuadScaling, then the
// final nodes from quad_CasimirPolder(n, alpha) as rpa_OptimizeFreqQuad does).
int grid_on_sets(int64_t nfreqs, int64_t nsets, const double *vals, const double *dw,
                 double *freqs, double *weights, double *info) {
    double alpha, rel, rmsd;
    if (!opt_scaling((int)nfreqs, nsets, vals, dw, alpha, rel, rmsd)) return BRPA_ERR_GRID;
    std::vector<double> x, w;
    if (!casimir_polder((int)nfreqs, alpha, x, w)) return BRPA_ERR_GRID;
    for (int64_t k = 0; k < nfreqs; ++k) {
        freqs[k] = x[k];
        weights[k] = w[k];
    }
    if (info) {
        info[0] = alpha;
        info[1] = rmsd;
        info[2] = rel;
    }
    return BRPA_OK;
}

int optimize_on_sets(int64_t nsets, const double *vals, const double *dw, double target_error,
                     double target_rel_error, int64_t min_nfreqs, int64_t max_nfreqs,
                     double *freqs, double *weights, int64_t *nfreqs, double *info) {
    for (int64_t n = min_nfreqs; n <= max_nfreqs; ++n) {
        double alpha, rel, rmsd;
        if (!opt_scaling((int)n, nsets, vals, dw, alpha, rel, rmsd)) return BRPA_ERR_GRID;
        if (rmsd < target_error && rel < target_rel_error) {
            std::vector<double> x, w;
            if (!casimir_polder((int)n, alpha, x, w)) return BRPA_ERR_GRID;
            for (int64_t k = 0; k < n; ++k) {
                freqs[k] = x[k];
                weights[k] = w[k];
            }
            *nfreqs = n;
            if (info) {
                info[0] = alpha;
                info[1] = rmsd;
                info[2] = rel;
            }
            return BRPA_OK;
        }
    }
    // "Quadrature did not converge to specified accuracy" (OptFreqQuad.f90:66-69)
    return BRPA_ERR_GRID;
}

bool sets_ok(int64_t nsets, const double *vals, const double *dw) {
    return nsets > 0 && nsets <= (int64_t)1 << 20 && vals && dw &&
           finite_all(vals, nsets * kNBins) && finite_all(dw, nsets * kNBins);
}

// the cap of a fragment's own histogram: no cap unless the caller gives one
double cap_of(double dai_max_thresh) {
    return dai_max_thresh > 0.0 ? dai_max_thresh : std::numeric_limits<double>::max();
}

}  // namespace

extern "C" {

int brpa_dai_histogram(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                       double core_orb_thresh, double dai_max_thresh, double *values,
                       double *weights) {
    if (!eocc || !evir || !values || !weights || nocc <= 0 || nvir <= 0 ||
        std::isnan(core_orb_thresh) || std::isnan(dai_max_thresh))
        return BRPA_ERR_ARG;
    if (!finite_all(eocc, nocc) || !finite_all(evir, nvir)) return BRPA_ERR_NONFINITE;
    return dai_histogram(eocc, nocc, evir, nvir, core_orb_thresh, cap_of(dai_max_thresh), values,
                         weights);
}

int brpa_dai_max_thresh(int64_t nspins, const double *eocc, const int64_t *nocc,
                        const double *evir, const int64_t *nvir, double core_orb_thresh,
                        double *dai_max_thresh) {
    if (nspins <= 0 || !eocc || !nocc || !evir || !nvir || !dai_max_thresh ||
        std::isnan(core_orb_thresh))
        return BRPA_ERR_ARG;
    double t = 0.0;
    int64_t oo = 0, ov = 0;
    for (int64_t s = 0; s < nspins; ++s) {
        if (nocc[s] < 0 || nvir[s] <= 0) return BRPA_ERR_ARG;
        if (!finite_all(eocc + oo, nocc[s]) || !finite_all(evir + ov, nvir[s]))
            return BRPA_ERR_NONFINITE;
        const int64_t nc = ncore(eocc + oo, nocc[s], core_orb_thresh);
        if (nc < nocc[s]) t = std::fmax(t, evir[ov + nvir[s] - 1] - eocc[oo + nc]);
        oo += nocc[s];
        ov += nvir[s];
    }
    *dai_max_thresh = t;
    return BRPA_OK;
}

int brpa_opt_quad_scaling(int64_t nfreqs, int64_t nsets, const double *values,
                          const double *dweights, double *freqs, double *weights, double *info) {
    if (!freqs || !weights || nfreqs < 2 || nfreqs > 440 || !sets_ok(nsets, values, dweights))
        return BRPA_ERR_ARG;
    return grid_on_sets(nfreqs, nsets, values, dweights, freqs, weights, info);
}

int brpa_optimize_freq_quad(int64_t nsets, const double *values, const double *dweights,
                            double target_error, double target_rel_error, int64_t min_nfreqs,
                            int64_t max_nfreqs, double *freqs, double *weights, int64_t *nfreqs,
                            double *info) {
    if (!freqs || !weights || !nfreqs || min_nfreqs < 2 || max_nfreqs < min_nfreqs ||
        max_nfreqs > 440 || !sets_ok(nsets, values, dweights))
        return BRPA_ERR_ARG;
    return optimize_on_sets(nsets, values, dweights, target_error, target_rel_error, min_nfreqs,
                            max_nfreqs, freqs, weights, nfreqs, info);
}

int brpa_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                   int64_t nfreqs, double *freqs, double *weights, double *info) {
    return brpa_freq_grid_ex(eocc, nocc, evir, nvir, nfreqs, BRPA_CORE_ORB_THRESH, 0.0, freqs,
                             weights, info);
}

int brpa_freq_grid_ex(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                      int64_t nfreqs, double core_orb_thresh, double dai_max_thresh,
                      double *freqs, double *weights, double *info) {
    if (!eocc || !evir || !freqs || !weights || nocc <= 0 || nvir <= 0 || nfreqs < 2 ||
        nfreqs > 440)
        return BRPA_ERR_ARG;
    double vals[kNBins], dw[kNBins];
    int rc = brpa_dai_histogram(eocc, nocc, evir, nvir, core_orb_thresh, dai_max_thresh, vals, dw);
    if (rc != BRPA_OK) return rc;
    if (ncore(eocc, nocc, core_orb_thresh) >= nocc) return BRPA_ERR_ARG;   // nothing active
    return grid_on_sets(nfreqs, 1, vals, dw, freqs, weights, info);
}

int brpa_optimize_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                            double target_error, double target_rel_error, int64_t min_nfreqs,
                            int64_t max_nfreqs, double *freqs, double *weights, int64_t *nfreqs,
                            double *info) {
    if (!eocc || !evir || !freqs || !weights || !nfreqs || nocc <= 0 || nvir <= 0 ||
        min_nfreqs < 2 || max_nfreqs < min_nfreqs || max_nfreqs > 440)
        return BRPA_ERR_ARG;
    double vals[kNBins], dw[kNBins];
    int rc = brpa_dai_histogram(eocc, nocc, evir, nvir, BRPA_CORE_ORB_THRESH, 0.0, vals, dw);
    if (rc != BRPA_OK) return rc;
    if (ncore(eocc, nocc, BRPA_CORE_ORB_THRESH) >= nocc) return BRPA_ERR_ARG;
    return optimize_on_sets(1, vals, dw, target_error, target_rel_error, min_nfreqs, max_nfreqs,
                            freqs, weights, nfreqs, info);
}

}  // extern "C"
