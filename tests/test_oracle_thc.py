"""Pins of the THC-route oracle (rpa_THC_PiU_BlockAlgo / rpa_THC_Rkab), CPU only."""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid, sosex, t2, thc


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


def _thc_amplitudes(o, v, n, g, seed):
    f = synth.make_thc_fragment(o, v, n, g, seed)
    Xa, Xi, Z = f.Xga.T, f.Xgi.T, f.Zgk.T
    R = thc.thc_rkai(Xa, Xi, Z)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    Pis = [erpa.gpiug_mo(R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(R, f.gaps(), x, w, Pis, seed=4)
    U = V.reshape(o, v, len(A)).transpose(1, 0, 2)      # U[a, i, mu]
    return f, Xa, Xi, Z, R, A, V, U


def test_thc_1b_is_twice_sosex():
    o, v = 3, 7
    f, Xa, Xi, Z, R, A, V, U = _thc_amplitudes(o, v, 12, 40, 2)
    Zgh = Z @ Z.T
    ec1b, ec2g, eig1b, eig2g, diag = thc.thc_1b2g(Zgh, Xi, Xa, U, A)
    assert abs(0.5 * ec1b - sosex.sosex(R, V, A, o, v)) < 1e-12 * abs(ec1b)
    assert abs(eig1b.sum() * 2.0 - ec1b) < 1e-12 * abs(ec1b)


def test_thc_2g_against_explicit_integrals_and_block_walk():
    o, v = 2, 5
    f, Xa, Xi, Z, R, A, V, U = _thc_amplitudes(o, v, 10, 30, 3)
    Zgh = Z @ Z.T
    ec1b, ec2g, _, eig2g, _ = thc.thc_1b2g(Zgh, Xi, Xa, U, A)
    assert abs(ec2g - thc.thc_1b2g_bruteforce_2g(Zgh, Xi, Xa, U, A)) < 1e-12 * abs(ec2g)
    b1, b2 = thc.thc_1b2g_blocked(Zgh, Xi, Xa, U, A, block=7)
    assert abs(b1 - ec1b) < 1e-12 * abs(ec1b) and abs(b2 - ec2g) < 1e-12 * abs(ec2g)
