"""Imaginary-frequency grid of the MO-basis RPA energy (oracle restatement).

Follows, step by step:

* ``rpa_NCore``          src/rpa/OrbDiffHist.f90:126-146
* ``rpa_DaiMaxThresh``   src/rpa/OrbDiffHist.f90:10-39
* ``rpa_DaiHistogram``   src/rpa/OrbDiffHist.f90:42-123
* ``quad_GaussLegendre`` src/rpa/quadratures.f90:418-506 (Newton in theta)
* ``quad_CasimirPolder`` src/rpa/quadratures.f90:570-643
* ``rpa_FreqQuadError``  src/rpa/OptFreqQuad.f90:125-169
* ``rpa_OptQuadScaling`` src/rpa/OptFreqQuad.f90:75-122
* ``rpa_OptimizeFreqQuad`` src/rpa/OptFreqQuad.f90:12-72

A grid is optimised on one or more histogram sets (one per spin and system,
src/rpa/rpa_driver.f90:118-172): the error of a scale alpha is the worst set's
(RMSD = sqrt(max over sets of sum_l w_l err_l^2), OptFreqQuad.f90:163-167).

Pinned by: Gauss-Legendre nodes/weights vs ``numpy.polynomial.legendre.leggauss``;
the Casimir-Polder rule integrating 1/(d^2+u^2) to pi/(2d) within the reported
RMSD; the histogram's bin/normalisation invariants and brute-force counts; core
freezing equal to dropping the core orbitals; the cap against a brute-force
count of the gaps below it; the all-core special values; the several-set error
against the pointwise maximum of single-set errors; brute-force re-scan of the
alpha grid (tests/test_oracle_grid.py).
"""
import math

import numpy as np

NBINS = 100                 # RPA_HISTOGRAM_NBINS, src/rpa/rpa_definitions.f90:113
MIN_NFREQS = 8              # RPA_MIN_NFREQS, src/rpa/rpa_definitions.f90:109
MAX_NFREQS = 64             # RPA_MAX_NFREQS, src/rpa/rpa_definitions.f90:110
CORE_ORB_THRESH = -3.0      # TRPAParams%CoreOrbThresh, src/rpa/rpa_definitions.f90:301
TARGET_ERROR_FREQ = 1.0e-6  # TRPAParams%TargetErrorFreq, src/rpa/rpa_definitions.f90:295
TARGET_REL_ERROR_FREQ = 1.0e-3  # src/rpa/rpa_definitions.f90:296
ALPHA_MIN, ALPHA_MAX, NSCAN = 0.1, 15.0, 15000  # src/rpa/OptFreqQuad.f90:92-94


def ncore(occ_energies, thresh=CORE_ORB_THRESH):
    """Number of leading occupied orbitals below ``thresh`` (OrbDiffHist.f90:126-146)."""
    n = 0
    for e in occ_energies:
        if e < thresh:
            n += 1
        else:
            break
    return n


def dai_max_thresh(spins, thresh=CORE_ORB_THRESH):
    """rpa_DaiMaxThresh (OrbDiffHist.f90:10-39): the largest active gap
    evir[-1] - eocc[ncore] over the spins ``[(eocc, evir), ...]`` of the whole
    complex; 0 when every spin's occupied orbitals are frozen."""
    t = 0.0
    for eocc, evir in spins:
        nc = ncore(eocc, thresh)
        if nc < len(eocc):
            t = max(t, float(evir[-1] - eocc[nc]))
    return t


def dai_histogram(eocc, evir, nbins=NBINS, dai_max_thresh=None, core_orb_thresh=CORE_ORB_THRESH):
    """100-bin histogram of orbital-energy gaps (OrbDiffHist.f90:42-123).

    ``eocc`` are ALL occupied energies of one spin (ascending, core included);
    the leading ones below ``core_orb_thresh`` are frozen (rpa_NCore, 69).
    ``evir`` are the virtual energies (ascending).  The range is
    [dmin, dmax] = [evir[0] - eocc[-1], min(cap, evir[-1] - eocc[ncore])]
    (73-75); bin k (0-based) has value dmin + k*w (its LEFT edge, 81-84); a gap
    goes to bin 1+nint((d-(dmin+w/2))/w) clamped to [1, nbins] (1-based,
    95-97); gaps above the cap are not counted (87).  ``dai_max_thresh`` None:
    no cap (the reference's huge(ONE)).  Returns (values, weights) with
    weights = count/total; with every occupied orbital frozen, the special
    values of 106-121 (weights -1, values[0] = huge, values[-1] = 0).
    """
    eocc = np.asarray(eocc, np.float64)
    evir = np.asarray(evir, np.float64)
    nc = ncore(eocc, core_orb_thresh)
    if nc >= eocc.size:
        values = np.zeros(nbins)
        values[0] = np.finfo(np.float64).max
        values[-1] = 0.0
        return values, -np.ones(nbins)
    cap = np.finfo(np.float64).max if dai_max_thresh is None else float(dai_max_thresh)
    dmin = evir[0] - eocc[-1]
    dmax = min(cap, evir[-1] - eocc[nc])
    if not dmin > 0.0:
        raise ValueError("Found invalid orbital energy gap. Cannot decompose energy denominators.")
    width = (dmax - dmin) / nbins
    values = dmin + np.arange(nbins, dtype=np.float64) * width
    d = (evir[None, :] - eocc[nc:, None]).ravel()
    d = d[d <= cap]
    # Fortran nint rounds half away from zero.  With every gap equal the width
    # is 0 and the quotient NaN: nint of it is the most negative integer on
    # x86 compilers, so max(1, k) puts the gap in the first bin.
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (d - (dmin + width / 2.0)) / width
    t = np.where(np.isnan(t), -1.0, np.clip(t, -2.0 * nbins, 2.0 * nbins))
    k = 1 + (np.sign(t) * np.floor(np.abs(t) + 0.5)).astype(np.int64)
    k = np.clip(k, 1, nbins)
    counts = np.bincount(k - 1, minlength=nbins)
    weights = np.abs(counts / float(counts.sum()))
    return values, weights


def _legendre(x, n):
    """P_n(x) and P_{n-1}(x) by the three-term recurrence."""
    p0 = np.ones_like(x)
    p1 = x.copy()
    if n == 0:
        return p0, np.zeros_like(x)
    for k in range(1, n):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
    return p1, p0


def gauss_legendre(n, tol=1.0e-14, maxit=20):
    """All n Gauss-Legendre nodes (ascending) and weights on [-1, 1].

    Newton iteration for the roots in the angle theta (x = cos theta), with the
    weight w = 2 / (dP_n/dtheta)^2, as in quad_GaussLegendre
    (quadratures.f90:418-506).  The initial guess is the Tricomi form
    theta_k = pi (4k-1)/(4n+2); the reference switches to the Gatteschi form
    near the interval ends, which changes only the starting point of a
    converged Newton iteration.
    """
    k = np.arange(1, n + 1, dtype=np.float64)
    theta = np.pi * (4.0 * k - 1.0) / (4.0 * n + 2.0)
    converged = False
    for _ in range(maxit):
        x = np.cos(theta)
        pa, pb = _legendre(x, n)
        dp = n * (x * pa - pb) / np.sin(theta)
        new = theta - pa / dp
        if np.all(np.abs(new - theta) < tol):
            theta = new
            converged = True
            break
        theta = new
    if not converged:
        raise RuntimeError("Gauss-Legendre Newton iteration did not converge for n=%d" % n)
    x = np.cos(theta)
    pa, pb = _legendre(x, n)
    dp = n * (x * pa - pb) / np.sin(theta)
    w = 2.0 / dp ** 2
    order = np.argsort(x)
    return x[order], w[order]


def casimir_polder(n, alpha=0.5):
    """Transformed GL rule on (0, inf): x = a(1+t)/(1-t), w = 2 w_GL a/(1-t)^2.

    quad_CasimirPolder, quadratures.f90:570-643.  Nodes ascending.
    """
    t, wgl = gauss_legendre(n)
    x = alpha * (1.0 + t) / (1.0 - t)
    w = 2.0 * wgl * alpha / (1.0 - t) ** 2
    return x, w


def _sets(values, dweights):
    """(nsets, nbins) views of one histogram (1-D) or several (rows)."""
    v = np.atleast_2d(np.asarray(values, np.float64))
    w = np.atleast_2d(np.asarray(dweights, np.float64))
    return v, w


def freq_quad_error(freqs, weights, values, dweights):
    """(max relative error, RMSD) of sum_k w_k/(d^2+x_k^2) vs pi/(2d).

    rpa_FreqQuadError, OptFreqQuad.f90:125-169: per set s, Err2_s = sum over
    bins with positive weight of w_l (approx - exact)^2; RMSD = sqrt(max_s
    Err2_s), the relative error is the largest over all sets.  ``values`` and
    ``dweights`` are one histogram or one per row.
    """
    V, Wt = _sets(values, dweights)
    rel, err2 = 0.0, 0.0
    for vals, dw_all in zip(V, Wt):
        m = dw_all > 0.0
        d = vals[m]
        dw = dw_all[m]
        if d.size == 0:
            continue
        approx = (weights[None, :] / (d[:, None] ** 2 + freqs[None, :] ** 2)).sum(axis=1)
        exact = np.pi / (2.0 * d)
        rel = max(rel, float(np.max(np.abs((approx - exact) / exact))))
        err2 = max(err2, float(np.sum(dw * (approx - exact) ** 2)))
    return rel, math.sqrt(err2)


def opt_quad_scaling(n, values, dweights):
    """Scan alpha_k = 0.1 + 14.9 (k-1)/14999 and keep the first strict RMSD minimum.

    rpa_OptQuadScaling, OptFreqQuad.f90:75-122, on one histogram or several
    (rows).  Returns (freqs, weights, alpha, max_rel_error, rmsd) with
    freqs = alpha * x_unscaled.
    """
    xu, wu = casimir_polder(n, 1.0)
    V, Wt = _sets(values, dweights)
    best_rmsd, best_rel, best_alpha = np.inf, 0.0, 0.0
    # Vectorised over all alphas, but every sum runs sequentially in the
    # reference's order (over nodes k, then over bins l): the RMSD minimum is
    # flat at the 1e-16 level, so the summation order decides which alpha is
    # the first strict minimum.  The selection rule is the sequential
    # "RMSD < OptRMSD" of the reference.
    ks = np.arange(NSCAN, dtype=np.float64)
    a = ALPHA_MIN + (ALPHA_MAX - ALPHA_MIN) * ks / float(NSCAN - 1)
    fx = a[:, None] * xu[None, :]
    fw = a[:, None] * wu[None, :]
    err2max = np.zeros(NSCAN)
    rel = np.zeros(NSCAN)
    for vals, dw_all in zip(V, Wt):
        m = dw_all > 0.0
        d = vals[m]
        dw = dw_all[m]
        exact = np.pi / (2.0 * d)
        err2 = np.zeros(NSCAN)
        for l in range(d.size):
            approx = np.zeros(NSCAN)
            for k in range(n):
                approx = approx + fw[:, k] / (d[l] * d[l] + fx[:, k] * fx[:, k])
            rel = np.fmax(rel, np.abs((approx - exact[l]) / exact[l]))
            err2 = err2 + dw[l] * (approx - exact[l]) * (approx - exact[l])
        err2max = np.fmax(err2max, err2)
    rmsd = np.sqrt(err2max)
    for j in range(NSCAN):
        if rmsd[j] < best_rmsd:
            best_rmsd, best_rel, best_alpha = float(rmsd[j]), float(rel[j]), float(a[j])
    return best_alpha * xu, best_alpha * wu, best_alpha, best_rel, best_rmsd


def freq_grid_sets(values, dweights, nfreqs):
    """Fixed-size grid on one or several histograms.

    The final nodes are regenerated by quad_CasimirPolder(n, alpha) as in
    rpa_OptimizeFreqQuad (OptFreqQuad.f90:60).  Returns (freqs, weights, info)
    with info = dict(alpha, rmsd, max_rel_error).
    """
    _, _, alpha, rel, rmsd = opt_quad_scaling(nfreqs, values, dweights)
    x, w = casimir_polder(nfreqs, alpha)
    return x, w, dict(alpha=alpha, rmsd=rmsd, max_rel_error=rel)


def freq_grid(eocc, evir, nfreqs, core_orb_thresh=CORE_ORB_THRESH, dai_max_thresh=None):
    """Fixed-size grid of one fragment on its own histogram (``eocc``: all
    occupied energies, core included)."""
    values, dweights = dai_histogram(eocc, evir, dai_max_thresh=dai_max_thresh,
                                     core_orb_thresh=core_orb_thresh)
    return freq_grid_sets(values, dweights, nfreqs)


def optimize_freq_quad(values, dweights, target_error=TARGET_ERROR_FREQ,
                       target_rel_error=TARGET_REL_ERROR_FREQ,
                       min_nfreqs=MIN_NFREQS, max_nfreqs=MAX_NFREQS):
    """Adaptive grid size on one or several histograms: the first n meeting both
    targets (rpa_OptimizeFreqQuad, OptFreqQuad.f90:12-72)."""
    for n in range(min_nfreqs, max_nfreqs + 1):
        _, _, alpha, rel, rmsd = opt_quad_scaling(n, values, dweights)
        if rmsd < target_error and rel < target_rel_error:
            x, w = casimir_polder(n, alpha)
            return x, w, dict(alpha=alpha, rmsd=rmsd, max_rel_error=rel, nfreqs=n)
    raise RuntimeError("Quadrature did not converge to specified accuracy")


def optimize_freq_grid(eocc, evir, target_error=TARGET_ERROR_FREQ,
                       target_rel_error=TARGET_REL_ERROR_FREQ,
                       min_nfreqs=MIN_NFREQS, max_nfreqs=MAX_NFREQS):
    """Adaptive grid of one fragment on its own histogram."""
    values, dweights = dai_histogram(eocc, evir)
    return optimize_freq_quad(values, dweights, target_error, target_rel_error, min_nfreqs,
                              max_nfreqs)
