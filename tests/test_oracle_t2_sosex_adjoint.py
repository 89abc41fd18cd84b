"""Pins of the amplitude, SOSEX and adjoint oracles (CPU only)."""
import numpy as np
import pytest

from brpa import synth
from oracle import adjoint, closed_forms as cf
from oracle import erpa, grid, sosex, t2


@pytest.fixture(scope="module")
def small():
    f = synth.make_fragment(2, 3, 4, seed=7)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 24)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    return f, x, w, Pis


def test_guess_vectors_distribution_and_determinism():
    a = t2.guess_vectors(123, 1000, 8)
    assert a.shape == (1000, 8) and np.all(a >= -0.5) and np.all(a < 0.5)
    assert abs(a.mean()) < 0.02 and abs(a.var() - 1 / 12) < 0.01
    assert np.array_equal(a, t2.guess_vectors(123, 1000, 8))
    assert not np.array_equal(a, t2.guess_vectors(124, 1000, 8))
    # column c of an (n, m) draw equals column c of an (n, m') draw (counter-based)
    assert np.array_equal(a[:, :3], t2.guess_vectors(123, 1000, 3))


def test_t2_full_space_equals_explicit(small):
    f, x, w, Pis = small
    I = t2.i_matrix_explicit(f.R, f.gaps(), x, w, Pis)
    Tex = t2.t2_explicit(I)
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=6, cutoff=0.0, seed=3)
    assert np.max(np.abs((V * A) @ V.T - Tex)) < 1e-11
    assert np.all(A <= 1e-14) and np.all(np.diff(A) >= 0)     # A ascending, T2 NSD
    assert np.allclose(V.T @ V, np.eye(V.shape[1]), atol=1e-7)   # S is ill-conditioned


def test_t2_matches_riccati_closed_form(small):
    """The frequency-integrated amplitudes equal 1/2 (1+K)^-1 (1-K) at the grid error."""
    f, _, _, _ = small
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 48)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    Tex = t2.t2_explicit(t2.i_matrix_explicit(f.R, f.gaps(), x, w, Pis))
    assert np.max(np.abs(Tex - cf.riccati_t2(f.R, f.gaps()))) < 1e-11


def test_eig_rpa_reproduces_erpa(small):
    """sum_mu 2 A_mu |R^T V_mu|^2 = 2 tr(V T2) = E_RPA (rpa_T2_EigRPA diagnostic)."""
    f, x, w, Pis = small
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=6, cutoff=0.0, seed=1)
    e_t2 = t2.eig_rpa(A, V, f.R).sum()
    assert abs(e_t2 - cf.plasmon_energy(f.R, f.gaps())) < 1e-8


def test_guess_count_is_not_capped():
    """ceiling(3 NAux) as rpa.f90:1696 passes it, also where it exceeds o*v."""
    assert t2.guess_count(351) == 1053
    assert t2.guess_count(351, 560) == 1053
    assert t2.guess_count(8, 36) == 24


def test_t2_oversampled_guess_space_is_seed_independent():
    """o=8, v=70, NAux=351, 16 frequencies (a shape of the 256-fragment batch):
    NG = 1053 > o*v = 560.  Every seed keeps all 560 directions, so E_x is the
    SOSEX of the explicit full-space amplitudes to 1e-9 Ha whatever the guess
    vectors.  A guess count capped at o*v (a square Gram matrix whose smallest
    eigenvalues fall below the cutoff) drops directions and misses that energy by
    more than 1e-5 Ha: the cap is a planted mistake this test catches."""
    o, v, n = 8, 70, 351
    f = synth.make_fragment(o, v, n, seed=11)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 16)
    g = f.gaps()
    Pis = [erpa.gpiug_mo(f.R, g, xx) for xx in x]
    a, E = np.linalg.eigh(t2.i_matrix_explicit(f.R, g, x, w, Pis))
    ex_full = sosex.sosex(f.R, E, 0.5 * a / (1.0 + a), o, v)
    worst_capped = 0.0
    for seed in (1, 2, 3):
        A, V = t2.t2_v5(f.R, g, x, w, Pis, seed=seed)
        assert len(A) == o * v
        assert abs(sosex.sosex(f.R, V, A, o, v) - ex_full) <= 1e-9
        Ac, Vc = t2.t2_v5(f.R, g, x, w, Pis, seed=seed, nguess=o * v)
        worst_capped = max(worst_capped, abs(sosex.sosex(f.R, Vc, Ac, o, v) - ex_full))
    assert worst_capped > 1e-5


def test_t2_randomized_subspace():
    """With NG < ov the range finder is approximate but E_T2 stays close to E_RPA."""
    f = synth.make_fragment(3, 12, 8, seed=2)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 16)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    ng = t2.guess_count(8, 36)
    assert ng == 24
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=5)
    e = erpa.ecrpa(f.R, f.gaps(), x, w)
    assert abs(t2.eig_rpa(A, V, f.R).sum() - e) < 1e-4 * abs(e)


def test_sosex_pair_loop_equals_bruteforce():
    f = synth.make_fragment(3, 4, 5, seed=11)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=12, cutoff=0.0, seed=2)
    e1 = sosex.sosex(f.R, V, A, 3, 4)
    e2 = sosex.sosex_bruteforce(f.R, V, A, 3, 4)
    assert abs(e1 - e2) < 1e-14 * max(1.0, abs(e2))
    assert abs(sosex.sosex(f.R, V, A, 3, 4, alpha=2.0) - 2 * e1) < 1e-15
    # the cumulant form rescales the amplitudes by 1/(1-4A^2)
    B = A / (1 - 4 * A ** 2)
    assert abs(sosex.sosex(f.R, V, A, 3, 4, s2_operator=True) - sosex.sosex(f.R, V, B, 3, 4)) < 1e-15


def test_sosex_planted_transpose_fails():
    """(ai|bj) instead of (aj|bi) (the direct term) is a different number."""
    f = synth.make_fragment(3, 4, 5, seed=11)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=12, cutoff=0.0, seed=2)
    T = (V * A) @ V.T
    direct = -np.sum((f.R @ f.R.T) * T)
    assert abs(direct - sosex.sosex(f.R, V, A, 3, 4)) > 1e-4


def test_adjoint_central_differences():
    f = synth.make_fragment(3, 6, 8, seed=13)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    dR, deo, dev = adjoint.ecrpa_grad(f.R, f.eocc, f.evir, x, w)
    h = 1e-5
    rng = np.random.default_rng(0)

    def E(R, eo, ev):
        return erpa.ecrpa(R, (ev[None] - eo[:, None]).ravel(), x, w, logdet=erpa.log_det_cholesky)

    for _ in range(8):
        i, j = rng.integers(0, 18), rng.integers(0, 8)
        Rp, Rm = f.R.copy(), f.R.copy()
        Rp[i, j] += h
        Rm[i, j] -= h
        fd = (E(Rp, f.eocc, f.evir) - E(Rm, f.eocc, f.evir)) / (2 * h)
        assert abs(fd - dR[i, j]) <= 1e-6 * max(abs(dR[i, j]), 1e-3)
    for k in range(3):
        ep, em = f.eocc.copy(), f.eocc.copy()
        ep[k] += h
        em[k] -= h
        fd = (E(f.R, ep, f.evir) - E(f.R, em, f.evir)) / (2 * h)
        assert abs(fd - deo[k]) <= 1e-6 * max(abs(deo[k]), 1e-3)
    for k in (0, 5):
        ep, em = f.evir.copy(), f.evir.copy()
        ep[k] += h
        em[k] -= h
        fd = (E(f.R, f.eocc, ep) - E(f.R, f.eocc, em)) / (2 * h)
        assert abs(fd - dev[k]) <= 1e-6 * max(abs(dev[k]), 1e-3)


def test_cancellation_free_fd_equals_plain_fd_on_small_energies():
    """On a tiny fragment (E ~ 0.1 Ha) and h = 1e-2, where the rounding of the
    plain FP64 difference (~1e-16 |E| / h) is below 1e-12, the cancellation-free
    central difference equals the plain one."""
    f = synth.make_fragment(2, 3, 4, seed=7)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    h = 1e-2
    fd = adjoint.central_fd_R(f.R, f.eocc, f.evir, x, w, [(3, 1)], h)[(3, 1)]
    Rp, Rm = f.R.copy(), f.R.copy()
    Rp[3, 1] += h
    Rm[3, 1] -= h
    plain = (erpa.ecrpa(Rp, f.gaps(), x, w, logdet=erpa.log_det_cholesky)
             - erpa.ecrpa(Rm, f.gaps(), x, w, logdet=erpa.log_det_cholesky)) / (2 * h)
    assert abs(fd - plain) < 1e-9 * abs(plain)
    v = adjoint.central_fd_eps(f.R, f.eocc, f.evir, x, w, "v", 1, h)
    ep, em = f.evir.copy(), f.evir.copy()
    ep[1] += h
    em[1] -= h
    plain = (erpa.ecrpa(f.R, (ep[None] - f.eocc[:, None]).ravel(), x, w)
             - erpa.ecrpa(f.R, (em[None] - f.eocc[:, None]).ravel(), x, w)) / (2 * h)
    assert abs(v - plain) < 1e-8 * abs(plain)
