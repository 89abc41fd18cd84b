"""E_RPA oracle pins: brute force, identities and closed forms (CPU only)."""
import math

import numpy as np
import pytest

from brpa import synth
from oracle import closed_forms as cf
from oracle import erpa, grid


@pytest.fixture(scope="module")
def tiny():
    f = synth.make_fragment(2, 3, 4, seed=7)
    return f, f.gaps()


def test_pi_bruteforce(tiny):
    f, g = tiny
    for u in (0.01, 0.5, 30.0):
        assert np.allclose(erpa.gpiug_mo(f.R, g, u), erpa.gpiug_mo_bruteforce(f.R, g, u),
                           rtol=1e-14, atol=1e-16)


def test_pi_symmetric_psd_and_prefactor():
    f = synth.make_fragment(3, 10, 12, seed=2)
    P = erpa.gpiug_mo(f.R, f.gaps(), 0.3)
    assert np.allclose(P, P.T, atol=1e-14)
    assert np.linalg.eigvalsh(P).min() > -1e-12
    assert np.allclose(erpa.gpiug_mo(f.R, f.gaps(), 0.3, 2.0), 0.5 * P, rtol=1e-15)
    acc = np.eye(12)
    erpa.gpiug_mo(f.R, f.gaps(), 0.3, Pi=acc)          # accumulate (beta = 1)
    assert np.allclose(acc, np.eye(12) + P, rtol=1e-15)


def test_sylvester_identity():
    """det(1_N + R^T D R) = det(1_ov + D^1/2 R R^T D^1/2): pins the scaling side."""
    f = synth.make_fragment(2, 4, 6, seed=5)
    g = f.gaps()
    u = 0.8
    P = erpa.gpiug_mo(f.R, g, u)
    D = 4 * g / (g ** 2 + u ** 2)
    M = np.eye(8) + np.sqrt(D)[:, None] * (f.R @ f.R.T) * np.sqrt(D)[None, :]
    assert abs(np.linalg.slogdet(np.eye(6) + P)[1] - np.linalg.slogdet(M)[1]) < 1e-12


def test_logdet_qr_equals_cholesky():
    f = synth.make_fragment(4, 20, 40, seed=1)
    P = erpa.gpiug_mo(f.R, f.gaps(), 0.2) + np.eye(40)
    assert abs(float(erpa.log_det_qr(P) - erpa.log_det_cholesky(P))) < 1e-11
    assert abs(float(erpa.log_det_qr(P)) - np.linalg.slogdet(P)[1]) < 1e-10


def test_plasmon_closed_form(tiny):
    """Quadrature E_RPA -> 1/2 tr[(D^1/2(D+4V)D^1/2)^1/2 - D - 2V] at the grid error."""
    f, g = tiny
    e_inf = cf.plasmon_energy(f.R, g)
    for n in (16, 32):
        x, w, info = grid.freq_grid(f.eocc, f.evir, n)
        e = erpa.ecrpa(f.R, g, x, w)
        assert abs(e - e_inf) <= max(1e-12, 100 * info["rmsd"]) * max(1.0, abs(e_inf))


def test_plasmon_closed_form_larger():
    f = synth.make_fragment(3, 12, 30, seed=4)
    g = f.gaps()
    x, w, info = grid.freq_grid(f.eocc, f.evir, 24)
    assert abs(erpa.ecrpa(f.R, g, x, w) - cf.plasmon_energy(f.R, g)) < 1e-9 * abs(
        cf.plasmon_energy(f.R, g))


def test_second_order_limit(tiny):
    """E(eps R)/eps^4 -> -2 sum V^2/(d+d') (direct MP2) with O(eps^2) corrections."""
    f, g = tiny
    e2 = cf.dmp2_energy(f.R, g)
    errs = []
    for eps in (0.1, 0.03, 0.01):
        errs.append(abs(cf.plasmon_energy(eps * f.R, g) / eps ** 4 - e2))
    assert errs[1] < errs[0] * 0.2 and errs[2] < errs[1] * 0.2
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 32)
    assert abs(erpa.ecrpa(0.03 * f.R, g, x, w) / 0.03 ** 4 - e2) < 5e-3 * abs(e2)


def test_lambda_integral_identity():
    """E_RPA = int_0^1 VcLambda dl (rpa_CC_VcLambdaPiU, rpa_CC.f90:341-361)."""
    f = synth.make_fragment(3, 8, 10, seed=9)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    assert abs(cf.ecrpa_lambda_integral(Pis, w) - erpa.ecrpa_analytic(Pis, x, w)) < 1e-12


def test_fused_equals_stored(tiny):
    f, g = tiny
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 10)
    Pis = [erpa.gpiug_mo(f.R, g, xx) for xx in x]
    e1, t1 = erpa.ecrpa(f.R, g, x, w, return_terms=True)
    e2, t2 = erpa.ecrpa_analytic(Pis, x, w, return_terms=True)
    assert e1 == e2 and t1 == t2


def test_planted_mistakes_are_caught(tiny):
    """A dropped prefactor, a wrong sign or a transposed product fails the closed form."""
    f, g = tiny
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 32)
    e_inf = cf.plasmon_energy(f.R, g)
    wrong = []
    wrong.append(erpa.ecrpa(f.R, g, x, w, prefactor=2.0))                 # dropped spin factor
    wrong.append(sum(ww / (2 * math.pi) * (np.linalg.slogdet(np.eye(4) - erpa.gpiug_mo(f.R, g, xx))[1]
                                            + np.trace(erpa.gpiug_mo(f.R, g, xx)))
                     for xx, ww in zip(x, w)))                            # wrong sign of Pi
    Rt = f.R.reshape(2, 3, 4).transpose(1, 0, 2).reshape(6, 4)            # ai order swapped
    wrong.append(erpa.ecrpa(Rt, g, x, w))
    for e in wrong:
        assert abs(e - e_inf) > 1e-6
