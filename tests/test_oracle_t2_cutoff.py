"""Pins of the oracle's amplitude-eigenvalue cutoff (rpa_THC_CC_T2,
rpa_CC_Doubles.f90:809-983) and smooth step (29-40), CPU only.

The pins are independent of the restatement: the step polynomial against the
integral it must be (a C^3 smooth step), the four threshold rules against
hand-worked answers on a fixed spectrum (including a value exactly at K, which
the rule rejects), the SUM-REJECTED rule against its definition as the
shortest kept prefix whose rejected tail sums below the threshold, and the
truncation against the Frobenius-norm identity of orthonormal eigenvectors.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from brpa import synth
from oracle import erpa, grid, t2

# ascending (largest |A| first), as the eigensolver returns them; -1e-4 sits exactly on K
SPEC = -np.array([0.5, 0.2, 0.05, 3e-3, 1e-4, 8e-5, 2e-5, 1e-6])


def _cut(A, *args, **kw):
    V = np.eye(len(A))
    return t2.t2_cutoff(A, V, *args, **kw)


def test_smooth_step_is_the_integrated_bump():
    """z(mu) = int_0^mu 35/16 (1 - t^2)^3 dt: alpha goes 1 -> 0 with three vanishing
    derivatives at both ends, through 1/2 at the midpoint."""
    ka, kb = 1e-4, 1e-5
    for mu in np.linspace(-1.0, 1.0, 17):
        lam = 0.5 * (mu * (kb - ka) + ka + kb)
        z = quad(lambda t: 35.0 / 16.0 * (1.0 - t * t) ** 3, 0.0, mu)[0]
        assert abs(t2.smooth_step(lam, ka, kb) - 0.5 * (1.0 - z)) < 1e-14
    assert t2.smooth_step(ka, ka, kb) == pytest.approx(1.0, abs=1e-15)
    assert t2.smooth_step(kb, ka, kb) == pytest.approx(0.0, abs=1e-15)
    assert t2.smooth_step(0.5 * (ka + kb), ka, kb) == pytest.approx(0.5, abs=1e-15)
    lam = np.linspace(kb, ka, 101)
    assert np.all(np.diff(t2.smooth_step(lam, ka, kb)) > 0)   # rises from kb to ka


@pytest.mark.parametrize("ctype,thresh,nocc,K,nkeep", [
    (t2.T2_CUTOFF_EIG, 1e-4, 1, 1e-4, 4),                 # 1e-4 <= K rejected
    (t2.T2_CUTOFF_EIG, 5e-5, 1, 5e-5, 6),
    (t2.T2_CUTOFF_EIG_DIV_MAXEIG, 1e-2, 1, 5e-3, 3),      # K = 1e-2 * 0.5
    (t2.T2_CUTOFF_EIG_DIV_NELECTRON, 1e-2, 5, 1e-3, 4),   # K = 1e-2 / 10
    # tails: after 5 (1-based) 1.01e-4, after 6 2.1e-5 < 1e-4 -> K = |A(7)| = 2e-5
    (t2.T2_CUTOFF_SUM_REJECTED, 1e-4, 1, 2e-5, 6),
    # no tail below 1e-9 except the empty one -> K = 0, everything kept
    (t2.T2_CUTOFF_SUM_REJECTED, 1e-9, 1, 0.0, 8),
])
def test_threshold_rules_known_answers(ctype, thresh, nocc, K, nkeep):
    A, V, Kk, common, nsc = _cut(SPEC, ctype, thresh, nocc)
    assert Kk == pytest.approx(K, rel=1e-14, abs=0.0)
    assert common == Kk and nsc == 0
    assert len(A) == nkeep and np.array_equal(A, SPEC[:nkeep])
    assert np.array_equal(V, np.eye(len(SPEC))[:, :nkeep])


def test_sum_rejected_is_shortest_prefix_with_small_tail():
    """Strictly decreasing |A|: NT2new = min{i >= 1 : |A(i+1) + ... + A(n)| < thresh}."""
    rng = np.random.default_rng(4)
    for trial in range(200):
        n = int(rng.integers(2, 40))
        A = -np.sort(10.0 ** rng.uniform(-9, 0, n))[::-1]
        thresh = 10.0 ** rng.uniform(-8, -1)
        tails = [abs(A[i:].sum()) for i in range(1, n)]
        want = next((i for i, s in enumerate(tails, start=1) if s < thresh), n)
        Ac, _, _, _, _ = _cut(A, t2.T2_CUTOFF_SUM_REJECTED, thresh, 1)
        assert len(Ac) == want, trial


def test_smooth_step_cut_known_answer():
    """K = 1e-4, Kb = 1e-5: 1e-4 kept whole (>= Ka), 8e-5 and 2e-5 scaled, 1e-6 cut."""
    A, _, K, _, nsc = _cut(SPEC, t2.T2_CUTOFF_EIG, 1e-4, 1, smooth_step_on=True, steepness=0.1)
    assert len(A) == 7 and nsc == 2 and K == 1e-4
    assert np.array_equal(A[:5], SPEC[:5])
    for i in (5, 6):
        assert A[i] == pytest.approx(SPEC[i] * t2.smooth_step(abs(SPEC[i]), 1e-4, 1e-5), rel=1e-15)
        assert abs(A[i]) < abs(SPEC[i])


def test_common_threshold_of_a_subsystem():
    """A subsystem applies the supersystem's K whatever its own rule would give."""
    _, _, K, common, _ = _cut(SPEC, t2.T2_CUTOFF_EIG_DIV_MAXEIG, 1e-2, 1)
    assert common == K == pytest.approx(5e-3)
    sub = SPEC * 0.05                          # its own MAXEIG rule would give 2.5e-4
    A, _, K2, common2, _ = _cut(sub, t2.T2_CUTOFF_EIG_DIV_MAXEIG, 1e-2, 1, common_thresh=common)
    assert K2 == common2 == common
    assert len(A) == 2                         # |-2.5e-3| <= 5e-3 ends the kept set


def test_no_vector_left_is_an_error():
    with pytest.raises(ValueError):
        _cut(SPEC, t2.T2_CUTOFF_EIG, 1.0, 1)
    with pytest.raises(ValueError):
        _cut(SPEC, 7, 1e-4, 1)


def test_truncated_amplitudes_frobenius_identity():
    """Orthonormal eigenvectors: |T2 - T2_cut|_F^2 = sum of the dropped A^2, and
    E_T2 changes by the dropped EigRPA terms."""
    f = synth.make_fragment(3, 8, 10, seed=11)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 16)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=24, cutoff=0.0, seed=9)
    K = np.sqrt(abs(A[3] * A[4]))              # between the 4th and 5th |A|
    Ac, Vc, _, _, _ = t2.t2_cutoff(A, V, t2.T2_CUTOFF_EIG, K)
    assert len(Ac) == 4
    d = (V * A) @ V.T - (Vc * Ac) @ Vc.T
    assert abs(np.sum(d * d) - np.sum(A[4:] ** 2)) < 1e-12
    e_full, e_cut = t2.eig_rpa(A, V, f.R), t2.eig_rpa(Ac, Vc, f.R)
    assert np.allclose(e_cut, e_full[:4], rtol=0, atol=1e-15)
