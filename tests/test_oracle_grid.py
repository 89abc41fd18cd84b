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


def _valence(seed, o=6, v=40):
    rng = np.random.default_rng(seed)
    return np.sort(rng.uniform(-1.5, -0.3, o)), np.geomspace(0.05, 3.0, v)


def test_core_freezing_equals_dropping_the_core():
    """Occupied orbitals below CoreOrbThresh = -3 Ha leave the histogram exactly
    as if they were absent (rpa_NCore, OrbDiffHist.f90:69-74); the library agrees."""
    import brpa
    eo, ev = _valence(1)
    full = np.concatenate([[-20.0, -10.0, -3.5], eo])
    v0, w0 = grid.dai_histogram(eo, ev)
    v1, w1 = grid.dai_histogram(full, ev)
    assert np.array_equal(v0, v1) and np.array_equal(w0, w1)
    assert grid.ncore(full) == 3
    # a different threshold freezes fewer orbitals and changes the histogram
    v2, w2 = grid.dai_histogram(full, ev, core_orb_thresh=-15.0)
    assert not np.array_equal(v2, v0)
    for thr in (grid.CORE_ORB_THRESH, -15.0):
        vl, wl = brpa.dai_histogram(full, ev, core_orb_thresh=thr)
        vr, wr = grid.dai_histogram(full, ev, core_orb_thresh=thr)
        assert np.array_equal(vl, vr) and np.array_equal(wl, wr)
    x, w, _ = brpa.freq_grid(full, ev, 16)
    xr, wr, _ = grid.freq_grid(eo, ev, 16)
    assert np.max(np.abs(x - xr) / xr) <= 1e-13 and np.max(np.abs(w - wr) / wr) <= 1e-13


def test_histogram_cap_against_brute_force():
    """A DaiMaxThresh below the fragment's own largest gap: the range ends at the
    cap and the gaps above it are not counted (OrbDiffHist.f90:75, 87)."""
    import brpa
    eo, ev = _valence(2)
    d = (ev[None, :] - eo[:, None]).ravel()
    cap = 0.5 * (d.min() + d.max())
    vals, wts = grid.dai_histogram(eo, ev, dai_max_thresh=cap)
    dmin = ev[0] - eo[-1]
    width = (cap - dmin) / grid.NBINS
    assert np.allclose(vals, dmin + width * np.arange(grid.NBINS), rtol=0, atol=1e-15)
    kept = d[d <= cap]
    cnt = np.zeros(grid.NBINS)
    for g in kept:                       # one gap at a time, the reference's loop
        t = (g - (dmin + width / 2)) / width
        k = 1 + int(math.copysign(math.floor(abs(t) + 0.5), t))
        cnt[min(max(k, 1), grid.NBINS) - 1] += 1
    assert kept.size < d.size and np.array_equal(wts, np.abs(cnt / kept.size))
    vl, wl = brpa.dai_histogram(eo, ev, dai_max_thresh=cap)
    assert np.array_equal(vl, vals) and np.array_equal(wl, wts)


def test_all_core_set_has_the_special_values_and_drops_out():
    """Every occupied orbital frozen: weights -1, values(1) = huge, values(NBins) = 0
    (OrbDiffHist.f90:106-121); such a set adds nothing to the grid error."""
    import brpa
    eo, ev = _valence(3)
    vs, ws = grid.dai_histogram([-9.0, -4.0], ev)
    assert np.all(ws == -1.0) and vs[0] == np.finfo(np.float64).max and vs[-1] == 0.0
    vl, wl = brpa.dai_histogram([-9.0, -4.0], ev)
    assert np.array_equal(vl, vs) and np.array_equal(wl, ws)
    v0, w0 = grid.dai_histogram(eo, ev)
    x, w = grid.casimir_polder(12, 0.9)
    assert grid.freq_quad_error(x, w, np.stack([v0, vs]), np.stack([w0, ws])) == \
        grid.freq_quad_error(x, w, v0, w0)


def test_dai_max_thresh_brute_force():
    """rpa_DaiMaxThresh: the largest active gap over the spins of the complex."""
    import brpa
    a_occ = np.array([-12.0, -1.2, -0.6])
    a_vir = np.array([0.1, 0.9, 2.5])
    b_occ = np.array([-1.4, -0.5])
    b_vir = np.array([0.2, 3.1])
    spins = [(a_occ, a_vir), (b_occ, b_vir)]
    want = max(2.5 - (-1.2), 3.1 - (-1.4))
    assert grid.dai_max_thresh(spins) == want
    assert brpa.dai_max_thresh(spins) == want
    assert grid.dai_max_thresh([(np.array([-5.0]), a_vir)]) == 0.0


def test_several_sets_take_the_worst_set():
    """Two histograms (say two monomers of a dimer): the grid error is the worse
    of the two at every alpha (OptFreqQuad.f90:163-167), the scan keeps a minimum
    of that maximum, and it differs from either set's own optimum, so a planted
    'first set only' rule fails; the library gives the same grid."""
    import brpa
    eo1, ev1 = _valence(4, o=5, v=30)
    eo2 = eo1 - 0.8
    ev2 = np.geomspace(0.3, 9.0, 30)
    h1, h2 = grid.dai_histogram(eo1, ev1), grid.dai_histogram(eo2, ev2)
    V, W = np.stack([h1[0], h2[0]]), np.stack([h1[1], h2[1]])
    x, w = grid.casimir_polder(10, 1.3)
    r1, m1 = grid.freq_quad_error(x, w, *h1)
    r2, m2 = grid.freq_quad_error(x, w, *h2)
    assert grid.freq_quad_error(x, w, V, W) == (max(r1, r2), max(m1, m2))
    xs, ws, info = grid.freq_grid_sets(V, W, 10)
    a = info["alpha"]
    step = (grid.ALPHA_MAX - grid.ALPHA_MIN) / (grid.NSCAN - 1)
    xu, wu = grid.casimir_polder(10, 1.0)
    for b in (a - step, a + step):
        assert grid.freq_quad_error(b * xu, b * wu, V, W)[1] >= info["rmsd"] * (1 - 1e-12)
    a1 = grid.freq_grid_sets(*h1, 10)[2]["alpha"]
    a2 = grid.freq_grid_sets(*h2, 10)[2]["alpha"]
    assert abs(a - a1) > 10 * step and abs(a - a2) > 10 * step
    xl, wl, il = brpa.freq_grid_sets(V, W, 10)
    assert np.max(np.abs(xl - xs) / xs) <= 1e-13 and np.max(np.abs(wl - ws) / ws) <= 1e-13
    xo, wo, io = grid.optimize_freq_quad(V, W)
    xl, wl, il = brpa.optimize_freq_quad(V, W)
    assert il["nfreqs"] == io["nfreqs"]
    assert np.allclose(xl, xo, rtol=1e-13) and np.allclose(wl, wo, rtol=1e-13)


def test_subsystem_grid_capped_at_the_complex():
    """GridLimitDai: the complex's largest gap caps its subsystems' histograms
    (rpa_driver.f90:123-147); one grid on [complex, A, B] from the library
    equals the oracle's."""
    import brpa
    eoa, eva = _valence(5, o=4, v=25)
    eob = np.sort(np.random.default_rng(6).uniform(-1.8, -0.4, 3))
    evb = np.geomspace(0.08, 6.0, 25)      # a ghost-basis monomer with high virtuals
    eoc = np.sort(np.concatenate([eoa, eob]))
    evc = np.sort(np.concatenate([eva, evb[:10]]))
    cap = grid.dai_max_thresh([(eoc, evc)])
    assert cap == brpa.dai_max_thresh([(eoc, evc)])
    hs = [grid.dai_histogram(e, v, dai_max_thresh=cap) for e, v in
          ((eoc, evc), (eoa, eva), (eob, evb))]
    hl = [brpa.dai_histogram(e, v, dai_max_thresh=cap) for e, v in
          ((eoc, evc), (eoa, eva), (eob, evb))]
    for (v0, w0), (v1, w1) in zip(hs, hl):
        assert np.array_equal(v0, v1) and np.array_equal(w0, w1)
    V, W = np.stack([h[0] for h in hs]), np.stack([h[1] for h in hs])
    x, w, _ = grid.freq_grid_sets(V, W, 16)
    xl, wl, _ = brpa.freq_grid_sets(V, W, 16)
    assert np.max(np.abs(xl - x) / x) <= 1e-13 and np.max(np.abs(wl - w) / w) <= 1e-13


@pytest.mark.parametrize("eo,ev", [
    ([-0.7], [0.4]),                      # one active pair
    ([-0.9, -0.7, -0.7], [0.4]),          # two distinct gaps (the ordinary rule)
    ([-12.0, -0.5, -0.5], [0.3, 0.3]),    # core frozen, every active gap equal
])
def test_degenerate_histogram_first_bin(eo, ev):
    """All active gaps equal (range of width 0): the bin quotient is 0/0 and the
    reference's clamp max(1, k) (OrbDiffHist.f90:95-97) puts every gap in the
    first bin; library and oracle agree, and the grid built on it integrates
    1/(d^2+u^2) at that one gap."""
    import brpa
    eo, ev = np.array(eo), np.array(ev)
    vals, wts = grid.dai_histogram(eo, ev)
    vl, wl = brpa.dai_histogram(eo, ev)
    assert np.array_equal(vl, vals) and np.array_equal(wl, wts)
    nc = grid.ncore(eo)
    d = (ev[None, :] - eo[nc:, None]).ravel()
    if np.ptp(d) == 0.0:
        assert wts[0] == 1.0 and np.all(wts[1:] == 0.0) and np.all(vals == d[0])
        x, w, _ = brpa.freq_grid(eo, ev, 8)
        xr, wr, _ = grid.freq_grid(eo, ev, 8)
        assert np.allclose(x, xr, rtol=1e-13) and np.allclose(w, wr, rtol=1e-13)
        g = d[0]
        assert abs(np.sum(w / (g * g + x * x)) * (2.0 * g / math.pi) - 1.0) <= 1e-6
    else:
        assert np.isclose(wts.sum(), 1.0)


def test_invalid_gap_and_arguments_are_errors():
    """A non-positive HOMO-LUMO gap stops the reference (OrbDiffHist.f90:76-79);
    the library returns an error code for it and for sizes it does not take."""
    import brpa
    eo, ev = _valence(5)
    with pytest.raises(ValueError):
        grid.dai_histogram([-0.5, 0.2], ev)
    with pytest.raises(brpa.BrpaError) as e:
        brpa.dai_histogram([-0.5, 0.2], ev)
    assert e.value.code == brpa._lib.ERR_GAP
    with pytest.raises(brpa.BrpaError) as e:
        brpa.dai_histogram(np.array([np.nan, -0.5]), ev)
    assert e.value.code == brpa._lib.ERR_NONFINITE
    for nf in (1, 441):
        with pytest.raises(brpa.BrpaError) as e:
            brpa.freq_grid(eo, ev, nf)
        assert e.value.code == brpa._lib.ERR_ARG
    with pytest.raises(brpa.BrpaError) as e:     # every occupied orbital frozen
        brpa.freq_grid(np.array([-9.0, -4.0]), ev, 8)
    assert e.value.code == brpa._lib.ERR_ARG
