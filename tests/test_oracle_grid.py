"""Frequency-grid oracle pins, and the library's host-side grid against it.

The reference holds no golden vectors for this path; these checks pin it:
Gauss-Legendre nodes/weights vs numpy's leggauss; the transformed rule
integrating 1/(d^2+u^2) to pi/(2d) within its reported RMSD; the histogram's
bin placement rule (OrbDiffHist.f90:81-97); brute-force re-scan of alpha.
"""
import math

import numpy as np
import pytest

from brpa import synth
from oracle import grid


@pytest.mark.parametrize("n", [2, 3, 8, 16, 24, 33, 64])
def test_gauss_legendre_matches_numpy(n):
    x, w = grid.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14
    assert np.max(np.abs(w - wr)) < 1e-14
    assert abs(w.sum() - 2.0) < 1e-14


def test_casimir_polder_transform():
    x, w = grid.casimir_polder(20, 0.7)
    t, wg = np.polynomial.legendre.leggauss(20)
    assert np.allclose(x, 0.7 * (1 + t) / (1 - t), rtol=1e-14)
    assert np.allclose(w, 2 * wg * 0.7 / (1 - t) ** 2, rtol=1e-14)
    assert np.all(np.diff(x) > 0)
    # integrates a rational test function exactly enough: int_0^inf 1/(1+u^2) = pi/2
    assert abs(np.sum(w / (1.0 + x ** 2)) - math.pi / 2) < 1e-6


def test_histogram_rules():
    eocc = np.array([-1.0, -0.5])
    evir = np.array([0.1, 0.2, 1.0])
    vals, wts = grid.dai_histogram(eocc, evir, nbins=4)
    dmin, dmax = 0.6, 2.0
    width = (dmax - dmin) / 4
    assert np.allclose(vals, dmin + width * np.arange(4))          # left edges
    d = np.sort((evir[None] - eocc[:, None]).ravel())
    # brute force of k = 1 + nint((d - (dmin + w/2))/w), clamped
    k = np.clip(1 + np.floor((d - (dmin + width / 2)) / width + 0.5).astype(int), 1, 4)
    cnt = np.bincount(k - 1, minlength=4)
    assert np.allclose(wts, cnt / cnt.sum())
    assert abs(wts.sum() - 1) < 1e-15
    with pytest.raises(ValueError):
        grid.dai_histogram(np.array([-1.0, 0.3]), evir)


def test_ncore():
    assert grid.ncore([-5.0, -3.5, -1.0, -0.2]) == 2
    assert grid.ncore([-1.0, -5.0]) == 0   # counting stops at the first active orbital


def test_optimized_grid_accuracy_and_scan():
    f = synth.make_fragment(8, 72, 400)
    x, w, info = grid.freq_grid(f.eocc, f.evir, 16)
    vals, dw = grid.dai_histogram(f.eocc, f.evir)
    rel, rmsd = grid.freq_quad_error(x, w, vals, dw)
    assert rmsd <= 1e-9 and abs(rmsd - info["rmsd"]) <= 1e-12
    m = dw > 0
    approx = (w[None] / (vals[m, None] ** 2 + x[None] ** 2)).sum(1)
    assert np.max(np.abs(approx - np.pi / (2 * vals[m]))) <= 10 * rmsd / math.sqrt(dw[m].min())
    # the chosen alpha is a minimum of the scan: its neighbours are not better
    step = (grid.ALPHA_MAX - grid.ALPHA_MIN) / (grid.NSCAN - 1)
    xu, wu = grid.casimir_polder(16, 1.0)
    for a in (info["alpha"] - step, info["alpha"] + step):
        _, r2 = grid.freq_quad_error(a * xu, a * wu, vals, dw)
        assert r2 >= info["rmsd"] * (1 - 1e-12)


def test_adaptive_grid_meets_targets():
    f = synth.make_fragment(4, 30, 100, seed=3)
    x, w, info = grid.optimize_freq_grid(f.eocc, f.evir)
    assert grid.MIN_NFREQS <= info["nfreqs"] <= grid.MAX_NFREQS
    assert info["rmsd"] < grid.TARGET_ERROR_FREQ and info["max_rel_error"] < grid.TARGET_REL_ERROR_FREQ


@pytest.mark.parametrize("o,v,n,nf", [(8, 72, 400, 16), (36, 300, 10, 24), (3, 9, 5, 8)])
def test_library_grid_matches_oracle(o, v, n, nf):
    """brpa_freq_grid is host code: it runs on any machine that loads the library."""
    import brpa
    f = synth.make_fragment(o, v, n)
    x, w, info = brpa.freq_grid(f.eocc, f.evir, nf)
    xr, wr, ir = grid.freq_grid(f.eocc, f.evir, nf)
    assert np.max(np.abs(x - xr) / xr) <= 1e-13
    assert np.max(np.abs(w - wr) / wr) <= 1e-13
    assert abs(info["alpha"] - ir["alpha"]) <= 2 * (grid.ALPHA_MAX - grid.ALPHA_MIN) / grid.NSCAN


def test_library_adaptive_grid_matches_oracle():
    import brpa
    f = synth.make_fragment(4, 30, 100, seed=3)
    x, w, info = brpa.optimize_freq_grid(f.eocc, f.evir)
    xr, wr, ir = grid.optimize_freq_grid(f.eocc, f.evir)
    assert info["nfreqs"] == ir["nfreqs"]
    assert np.allclose(x, xr, rtol=1e-13) and np.allclose(w, wr, rtol=1e-13)
