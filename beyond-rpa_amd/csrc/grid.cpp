// Imaginary-frequency grid of the MO-basis RPA energy (host code).
//
//   dai_histogram    rpa_DaiHistogram   rpa/OrbDiffHist.f90:42-123
//   gauss_legendre   quad_GaussLegendre rpa/quadratures.f90:418-506 (Newton in theta)
//   casimir_polder   quad_CasimirPolder rpa/quadratures.f90:570-643
//   quad_error       rpa_FreqQuadError  rpa/OptFreqQuad.f90:125-169
//   opt_scaling      rpa_OptQuadScaling rpa/OptFreqQuad.f90:75-122
//   brpa_optimize_freq_grid  rpa_OptimizeFreqQuad rpa/OptFreqQuad.f90:12-72
#include <cmath>
#include <limits>
#include <vector>

#include "internal.h"

namespace {

constexpr int kNBins = 100;            // RPA_HISTOGRAM_NBINS (rpa_definitions.f90:113)
constexpr double kAlphaMin = 0.1, kAlphaMax = 15.0;
constexpr int kNScan = 15000;          // OptFreqQuad.f90:92-94
constexpr double kPi = 3.14159265358979323846;

int dai_histogram(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                  std::vector<double> &values, std::vector<double> &weights) {
    double dmin = evir[0] - eocc[nocc - 1];
    double dmax = evir[nvir - 1] - eocc[0];
    const double cap = dmax;   // one fragment: the cap is its own largest gap
    if (!(dmin > 0.0)) return BRPA_ERR_GAP;
    const double width = (dmax - dmin) / kNBins;
    values.resize(kNBins);
    weights.assign(kNBins, 0.0);
    for (int k = 0; k < kNBins; ++k) values[k] = dmin + k * width;
    std::vector<int64_t> count(kNBins, 0);
    int64_t total = 0;
    for (int64_t i = 0; i < nocc; ++i)
        for (int64_t a = 0; a < nvir; ++a) {
            double d = evir[a] - eocc[i];
            if (d <= cap) {
                double t = (d - (dmin + width / 2.0)) / width;
                // Fortran nint: round half away from zero
                int64_t k = 1 + (int64_t)(t >= 0 ? std::floor(t + 0.5) : -std::floor(-t + 0.5));
                if (k < 1) k = 1;
                if (k > kNBins) k = kNBins;
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

void quad_error(const double *fx, const double *fw, int n, const std::vector<double> &vals,
                const std::vector<double> &dw, double &rel, double &rmsd) {
    double err2 = 0.0;
    rel = 0.0;
    for (size_t l = 0; l < vals.size(); ++l) {
        if (dw[l] > 0.0) {
            const double d = vals[l];
            const double exact = kPi / (2.0 * d);
            double approx = 0.0;
            for (int k = 0; k < n; ++k) approx += fw[k] / (d * d + fx[k] * fx[k]);
            rel = std::fmax(rel, std::fabs((approx - exact) / exact));
            err2 += dw[l] * (approx - exact) * (approx - exact);
        }
    }
    rmsd = std::sqrt(err2);
}

bool opt_scaling(int n, const std::vector<double> &vals, const std::vector<double> &dw,
                 double &alpha, double &rel, double &rmsd) {
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
        quad_error(fx.data(), fw.data(), n, vals, dw, r, m);
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

}  // namespace

extern "C" {

int brpa_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                   int64_t nfreqs, double *freqs, double *weights, double *info) {
    if (!eocc || !evir || !freqs || !weights || nocc <= 0 || nvir <= 0 || nfreqs < 2 ||
        nfreqs > 440)
        return BRPA_ERR_ARG;
    if (!finite_all(eocc, nocc) || !finite_all(evir, nvir)) return BRPA_ERR_NONFINITE;
    std::vector<double> vals, dw;
    int rc = dai_histogram(eocc, nocc, evir, nvir, vals, dw);
    if (rc != BRPA_OK) return rc;
    double alpha, rel, rmsd;
    if (!opt_scaling((int)nfreqs, vals, dw, alpha, rel, rmsd)) return BRPA_ERR_GRID;
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

int brpa_optimize_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                            double target_error, double target_rel_error, int64_t min_nfreqs,
                            int64_t max_nfreqs, double *freqs, double *weights, int64_t *nfreqs,
                            double *info) {
    if (!eocc || !evir || !freqs || !weights || !nfreqs || nocc <= 0 || nvir <= 0 ||
        min_nfreqs < 2 || max_nfreqs < min_nfreqs || max_nfreqs > 440)
        return BRPA_ERR_ARG;
    if (!finite_all(eocc, nocc) || !finite_all(evir, nvir)) return BRPA_ERR_NONFINITE;
    std::vector<double> vals, dw;
    int rc = dai_histogram(eocc, nocc, evir, nvir, vals, dw);
    if (rc != BRPA_OK) return rc;
    for (int64_t n = min_nfreqs; n <= max_nfreqs; ++n) {
        double alpha, rel, rmsd;
        if (!opt_scaling((int)n, vals, dw, alpha, rel, rmsd)) return BRPA_ERR_GRID;
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
    return BRPA_ERR_GRID;
}

}  // extern "C"
