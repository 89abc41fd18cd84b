"""Pins of the THC-route oracle (rpa_THC_PiU_BlockAlgo / rpa_THC_Rkab), CPU only."""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, thc


def _f(o, v, n, g, seed):
    f = synth.make_thc_fragment(o, v, n, g, seed)
    return f, f.Xga.T, f.Xgi.T, f.Zgk.T


@pytest.mark.parametrize("o,v,n,g", [(1, 1, 1, 1), (2, 3, 4, 5), (3, 5, 7, 11)])
def test_thc_rkai_equals_triple_loop(o, v, n, g):
    _, Xa, Xi, Z = _f(o, v, n, g, 3)
    R = thc.thc_rkai(Xa, Xi, Z)
    assert R.shape == (o * v, n)
    assert np.max(np.abs(R - thc.thc_rkai_bruteforce(Xa, Xi, Z))) < 1e-15


def test_thc_rkai_per_occupied_association():
    """rpa_THC_Rkai's order (one occupied at a time) gives the same R."""
    _, Xa, Xi, Z = _f(4, 9, 13, 40, 5)
    R = thc.thc_rkai(Xa, Xi, Z)
    assert np.max(np.abs(R - thc.thc_rkai_per_occupied(Xa, Xi, Z))) < 1e-14 * np.max(np.abs(R))


def test_planted_index_order_fails():
    """ai with i fastest (a wrong index map) is caught by the triple loop."""
    _, Xa, Xi, Z = _f(2, 3, 4, 5, 7)
    R = thc.thc_rkai_bruteforce(Xa, Xi, Z)
    G = Xa.shape[0]
    wrong = (Xa[:, :, None] * Xi[:, None, :]).reshape(G, -1).T @ Z   # i fastest
    assert np.max(np.abs(wrong - R)) > 1e-3


def test_thc_piu_properties():
    f, Xa, Xi, Z = _f(3, 8, 10, 30, 9)
    freqs = np.array([0.05, 0.7, 3.0])
    Pis, R = thc.thc_piu(Xa, Xi, Z, f.eocc, f.evir, freqs)
    for k, u in enumerate(freqs):
        P = Pis[k]
        assert np.array_equal(P, erpa.gpiug_mo(R, f.gaps(), u))
        assert np.max(np.abs(P - P.T)) < 1e-15
        assert np.linalg.eigvalsh(P).min() > -1e-14
        assert np.max(np.abs(P - erpa.gpiug_mo_bruteforce(R, f.gaps(), u))) < 1e-13
    # closed shell (4) = two open-shell spin blocks (2 each) of the same orbitals
    P2, _ = thc.thc_piu(Xa, Xi, Z, f.eocc, f.evir, freqs, prefactor=2.0)
    assert np.max(np.abs(Pis - 2.0 * P2)) < 1e-15
