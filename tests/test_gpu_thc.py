"""GPU parity of the THC route's R producer and Pi(u) (brpa_thc_piu) against
the CPU oracle (oracle/thc.py).

Tolerances (FP64 on both sides, different summation orders): R within 1e-13
of its largest element, Pi within 1e-12 of its largest element, E_RPA from
those Pi within 1e-8 Ha.
"""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid, thc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("o,v,n,g,nf", [(1, 1, 1, 1, 1), (2, 3, 5, 7, 3), (3, 21, 70, 37, 4),
                                         (5, 60, 200, 333, 8), (12, 90, 500, 1500, 16)])
def test_thc_piu_matches_oracle(ctx, o, v, n, g, nf):
    """Ragged sizes: ov, naux not multiples of the 64-wide tiles, ngrid not a
    multiple of the 16-point stage, fewer frequencies than the 4 per work-group."""
    f = synth.make_thc_fragment(o, v, n, g, 100 + o)
    freqs = np.geomspace(0.02, 20.0, nf)
    Pis_ref, R_ref = thc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, freqs)
    Pi, R = ctx.thc_piu(f.Xga, f.Xgi, f.Zgk, f.eocc, f.evir, freqs, want_R=True)
    print("R rel", _rel(R, R_ref), "Pi rel", _rel(Pi, Pis_ref))
    assert _rel(R, R_ref) <= 1e-13
    assert _rel(Pi, Pis_ref) <= 1e-12


def test_thc_piu_open_shell_accumulates(ctx):
    """Two spin blocks (prefactor 2 each) summed into one PiU, as rpa_THC_PiU
    does for NSpins = 2 (rpa_THC.f90:262-281)."""
    fa = synth.make_thc_fragment(4, 30, 90, 200, 1)
    fb = synth.make_thc_fragment(3, 31, 90, 200, 2)
    fb.Zgk = fa.Zgk                       # Zgk is shared between the spins
    freqs = np.array([0.1, 1.0, 5.0])
    ref = sum(thc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, freqs, 2.0)[0]
              for f in (fa, fb))
    Pi = ctx.thc_piu(fa.Xga, fa.Xgi, fa.Zgk, fa.eocc, fa.evir, freqs, prefactor=2.0)
    Pi = ctx.thc_piu(fb.Xga, fb.Xgi, fb.Zgk, fb.eocc, fb.evir, freqs, prefactor=2.0, PiU=Pi)
    assert _rel(Pi, ref) <= 1e-12


def test_thc_piu_device_tensors_and_ecrpa(ctx):
    """torch tensors in HBM end to end; E_RPA of the THC Pi (brpa_ecrpa_analytic)
    and of the produced R (brpa_ecrpa) equal the oracle's E_RPA."""
    import torch
    o, v, n, g, nf = 8, 72, 400, 900, 16
    f = synth.make_thc_fragment(o, v, n, g, 77)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis_ref, R_ref = thc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, x)
    e_ref = erpa.ecrpa_analytic(Pis_ref, x, w)
    d = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    Pi, R = ctx.thc_piu(d(f.Xga), d(f.Xgi), d(f.Zgk), f.eocc, f.evir, x, want_R=True)
    assert Pi.is_cuda and R.is_cuda
    st = ctx.last_thc_stage_ms()
    assert st["thc_r"][1] == 1 and st["thc_pi"][1] == 1
    assert st["thc_r"][0] > 0.0 and st["thc_pi"][0] > 0.0
    assert _rel(R.cpu().numpy(), R_ref) <= 1e-13
    e1 = ctx.ecrpa_analytic(Pi, x, w)[0]
    e2 = ctx.ecrpa(R, f.eocc, f.evir, x, w)[0]
    print("E_RPA", e_ref, e1 - e_ref, e2 - e_ref)
    assert abs(e1 - e_ref) <= 1e-8 and abs(e2 - e_ref) <= 1e-8


def test_thc_piu_empty_spin_block(ctx):
    """nocc = 0: nothing to add (accumulate) or a zero PiU (overwrite)."""
    f = synth.make_thc_fragment(2, 5, 16, 20, 3)
    freqs = np.array([0.5, 2.0])
    Pi = ctx.thc_piu(f.Xga, f.Xgi, f.Zgk, f.eocc, f.evir, freqs)
    before = Pi.copy()
    empty = np.zeros((0, f.ngrid))
    Pi = ctx.thc_piu(f.Xga, empty, f.Zgk, np.zeros(0), f.evir, freqs, PiU=Pi)
    assert np.array_equal(Pi, before)
    Z = ctx.thc_piu(f.Xga, empty, f.Zgk, np.zeros(0), f.evir, freqs)
    assert not np.any(Z)


@pytest.mark.parametrize("o,v,n,g", [(2, 5, 10, 30), (3, 21, 70, 137), (6, 40, 120, 333)])
def test_thc_1b2g_matches_oracle(ctx, o, v, n, g):
    """THC SOSEX + 2g (rpa_CCD_corrections_1b2g): energies and per-vector terms
    within 1e-11 relative of the oracle, YXUggm within 1e-12; Ec1b / 2 equals the
    library's own SOSEX of the same amplitudes on the THC-built R."""
    from oracle import t2
    f = synth.make_thc_fragment(o, v, n, g, 9)
    Xa, Xi, Z = f.Xga.T, f.Xgi.T, f.Zgk.T
    R = thc.thc_rkai(Xa, Xi, Z)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 10)
    Pis = [erpa.gpiug_mo(R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(R, f.gaps(), x, w, Pis, seed=6)
    U = V.reshape(o, v, len(A)).transpose(1, 0, 2)
    Zgh = Z @ Z.T
    ref = thc.thc_1b2g(Zgh, Xi, Xa, U, A)
    Uaim = np.ascontiguousarray(V.T)                 # Fortran Uaim(v, o, NT2)
    e1, e2, es, eg, diag = ctx.thc_1b2g(np.ascontiguousarray(Zgh), f.Xgi, f.Xga, Uaim, A,
                                        want_diag=True)
    print("NT2", len(A), "1b", ref[0], e1 - ref[0], "2g", ref[1], e2 - ref[1])
    assert abs(e1 - ref[0]) <= 1e-11 * abs(ref[0])
    assert abs(e2 - ref[1]) <= 1e-11 * abs(ref[1])
    assert np.max(np.abs(es - ref[2])) <= 1e-11 * np.max(np.abs(ref[2]))
    assert np.max(np.abs(eg - ref[3])) <= 1e-11 * np.max(np.abs(ref[3]))
    assert np.max(np.abs(diag.T - ref[4])) <= 1e-12 * np.max(np.abs(ref[4]))
    ex = ctx.sosex(V, A, R, o, v)
    assert abs(0.5 * e1 - ex) <= 1e-10 * abs(ex)
