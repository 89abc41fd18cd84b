"""Pins of the MBPT3-1 / MBPT3-2 exchange oracle (oracle/mbpt3.py), CPU only."""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid, mbpt3, sosex, t2


@pytest.fixture(scope="module")
def full_space():
    """o=2, v=3, N=4, NG = ov, zero cutoff, 48 frequencies."""
    f = synth.make_fragment(2, 3, 4, 7)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 48)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=6, cutoff=0.0, seed=3)
    return f, x, w, Pis, A, V


def test_mbpt3_1_is_the_coupling_derivative_of_sosex(full_space):
    """E_MBPT3_1 = d/dlambda E_SOSEX(T(lambda)) at lambda = 1: the reference's
    numerical variant (rpa.f90:1589-1604, 5-point stencil) on the same grid."""
    f, x, w, Pis, A, V = full_space
    e1 = mbpt3.mbpt3_1(f.R, V, A, 2, 3, f.eocc, f.evir)
    h = 1e-3

    def ex(lam):
        A_, V_ = t2.t2_v5(f.R, f.gaps(), x, w, Pis, lam=lam, nguess=6, cutoff=0.0, seed=3)
        return sosex.sosex(f.R, V_, A_, 2, 3)
    fd = (ex(1 - 2 * h) - 8 * ex(1 - h) + 8 * ex(1 + h) - ex(1 + 2 * h)) / (12 * h)
    assert abs(e1 - fd) < 1e-10 * abs(fd)


def test_planted_sign_in_bmat_fails(full_space):
    """A flipped 4 U A term (a plausible slip) breaks the derivative identity."""
    f, x, w, Pis, A, V = full_space
    Am, Bm = mbpt3.mbpt3_1_matrices(f.R, V, A, f.eocc, f.evir)
    W = f.R.T @ V
    U = W.T @ W
    wrong = mbpt3.lyapunov(Bm - 8.0 * U * A[None, :], Am)
    right = mbpt3.lyapunov(Bm, Am)
    e_w = -mbpt3.trxk(f.R, V, wrong, 2, 3)
    e_r = -mbpt3.trxk(f.R, V, right, 2, 3)
    assert abs(e_w - e_r) > 1e-6 * abs(e_r)


def test_lyapunov_against_kronecker_solve():
    rng = np.random.default_rng(2)
    n = 7
    M = rng.standard_normal((n, n))
    B = M @ M.T + n * np.eye(n) + 0.3 * rng.standard_normal((n, n))   # nonsymmetric, stable
    S = rng.standard_normal((n, n))
    Am = S + S.T
    X = mbpt3.lyapunov(B, Am)
    # vec(B^T X + X B) = (I kron B^T + B^T kron I) vec(X), column-major vec
    L = np.kron(np.eye(n), B.T) + np.kron(B.T, np.eye(n))
    Xk = np.linalg.solve(L, Am.ravel(order="F")).reshape((n, n), order="F")
    assert np.max(np.abs(X - Xk)) < 1e-13 * np.max(np.abs(Xk))
    Xs, it = mbpt3.lyapunov_sign(B, Am)
    assert np.max(np.abs(Xs - Xk)) < 1e-12 * np.max(np.abs(Xk)) and it < 30


def test_trxk_and_mbpt3_2_against_explicit_tensors():
    f = synth.make_fragment(3, 4, 6, 5)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=1)
    rng = np.random.default_rng(3)
    S = rng.standard_normal((len(A), len(A)))
    X = S + S.T
    assert abs(mbpt3.trxk(f.R, V, X, 3, 4) - mbpt3.trxk_bruteforce(f.R, V, X, 3, 4)) < 1e-13
    e2 = mbpt3.mbpt3_2(f.R, V, A, 3, 4)
    assert abs(e2 - mbpt3.mbpt3_2_bruteforce(f.R, V, A, 3, 4)) < 1e-13 * abs(e2)


def test_mbpt3_2_scales_as_sixth_power():
    """T is second order in R at small R, so E_MBPT3_2(eps R) / eps^6 tends to a
    constant (a third-order energy)."""
    f = synth.make_fragment(2, 5, 6, 9)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 32)
    vals = []
    for eps in (1e-2, 5e-3):
        R = eps * f.R
        Pis = [erpa.gpiug_mo(R, f.gaps(), xx) for xx in x]
        A, V = t2.t2_v5(R, f.gaps(), x, w, Pis, nguess=10, cutoff=0.0, seed=1)
        vals.append(mbpt3.mbpt3_2(R, V, A, 2, 5) / eps ** 6)
    assert abs(vals[0] - vals[1]) < 1e-3 * abs(vals[1])
