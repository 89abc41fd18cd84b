"""GPU parity of the amplitude (T2) and SOSEX paths against the CPU oracle.

The guess vectors are the same bits on both sides (counter hash), so the
randomised range finder sees the same subspace; the comparisons use
basis-independent quantities (the T2 matrix, its eigenvalues, E_x).
Tolerances: 1e-8 Ha on exchange energies (FP64), 1e-9 relative on T2.
"""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid, sosex, t2

pytestmark = pytest.mark.gpu


def _setup(o, v, n, nf, seed):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    return f, x, w, Pis


@pytest.mark.parametrize("o,v,n,nf,ng", [(2, 3, 4, 12, 6), (3, 10, 12, 10, 0), (4, 25, 40, 12, 0),
                                         (6, 40, 90, 16, 0)])
def test_t2_matches_oracle(ctx, o, v, n, nf, ng):
    f, x, w, Pis = _setup(o, v, n, nf, 31)
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, nguess=ng or None, seed=77)
    A, V = ctx.t2(f.R, f.eocc, f.evir, x, w, nguess=ng, seed=77)
    print("NT2", len(A), len(A_ref))
    assert len(A) == len(A_ref)
    assert np.max(np.abs(A - A_ref)) <= 1e-9 * np.max(np.abs(A_ref))
    T_ref = (V_ref * A_ref) @ V_ref.T
    T = (V * A) @ V.T
    assert np.max(np.abs(T - T_ref)) <= 1e-9 * np.max(np.abs(T_ref))
    # columns tied to Gram eigenvalues near the cutoff are only as orthonormal as
    # s^(-1/2) amplification allows, in the reference's algorithm as well
    dev_ref = np.max(np.abs(V_ref.T @ V_ref - np.eye(len(A_ref))))
    assert np.max(np.abs(V.T @ V - np.eye(len(A)))) <= 100 * dev_ref + 1e-9


def test_t2_full_space_invariants(ctx):
    """NG = ov and a zero cutoff: T2 = I(1+I)^-1/2 and E_T2 = E_RPA."""
    f, x, w, Pis = _setup(2, 3, 4, 24, 7)
    A, V = ctx.t2(f.R, f.eocc, f.evir, x, w, nguess=6, cutoff=0.0, seed=3)
    Tex = t2.t2_explicit(t2.i_matrix_explicit(f.R, f.gaps(), x, w, Pis))
    assert np.max(np.abs((V * A) @ V.T - Tex)) < 1e-11
    assert abs(t2.eig_rpa(A, V, f.R).sum() - erpa.ecrpa(f.R, f.gaps(), x, w)) < 1e-10


@pytest.mark.parametrize("o,v,n", [(3, 4, 5), (5, 20, 30), (7, 33, 80)])
def test_sosex_matches_oracle(ctx, o, v, n):
    f, x, w, Pis = _setup(o, v, n, 10, 41)
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=5)
    ex_ref = sosex.sosex(f.R, V, A, o, v)
    ex = ctx.sosex(V, A, f.R, o, v)
    print("Ex", ex, ex_ref, ex - ex_ref)
    assert abs(ex - ex_ref) <= 1e-8 * max(1.0, abs(ex_ref)) and abs(ex - ex_ref) <= 1e-8
    assert abs(ctx.sosex(V, A, f.R, o, v, alpha=2.0) - 2 * ex) <= 1e-12 * max(1, abs(ex))
    ex_s2 = ctx.sosex(V, A, f.R, o, v, s2_operator=True)
    assert abs(ex_s2 - sosex.sosex(f.R, V, A, o, v, s2_operator=True)) <= 1e-8


def test_rpt2_chain_matches_oracle(ctx):
    """GPU T2 -> GPU SOSEX equals oracle T2 -> oracle SOSEX on the same guess vectors."""
    o, v, n = 6, 30, 70
    f, x, w, Pis = _setup(o, v, n, 12, 12)
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=2026)
    ex_ref = sosex.sosex(f.R, V_ref, A_ref, o, v)
    A, V = ctx.t2(f.R, f.eocc, f.evir, x, w, seed=2026)
    ex = ctx.sosex(V, A, f.R, o, v)
    print("chain Ex", ex, ex_ref, ex - ex_ref)
    assert abs(ex - ex_ref) <= 1e-8
