"""GPU parity of the third-order exchange corrections MBPT3-1 and MBPT3-2
(brpa_mbpt3_1 / brpa_mbpt3_2) against the CPU oracle (oracle/mbpt3.py).

The amplitudes come from the oracle (same A, V on both sides).  Tolerances:
FP64 1e-10 relative to |E| (the library solves the Lyapunov equation by a
sign iteration, the oracle by Bartels-Stewart); the split-int8 final
contraction of MBPT3-1 within 1e-7 Ha.
"""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid, mbpt3, t2

pytestmark = pytest.mark.gpu


def _amps(o, v, n, nf, seed, **kw):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=77, **kw)
    return f, x, w, A, V


@pytest.mark.parametrize("o,v,n,nf,ng", [(2, 3, 4, 24, 6), (3, 10, 12, 10, 0), (6, 40, 90, 12, 0),
                                         (12, 60, 324, 12, 0)])
def test_mbpt3_matches_oracle(ctx, o, v, n, nf, ng):
    kw = {"nguess": ng, "cutoff": 0.0} if ng else {}
    f, x, w, A, V = _amps(o, v, n, nf, 31, **kw)
    e1_ref = mbpt3.mbpt3_1(f.R, V, A, o, v, f.eocc, f.evir)
    e2_ref = mbpt3.mbpt3_2(f.R, V, A, o, v)
    e1 = ctx.mbpt3_1(V, A, f.R, f.eocc, f.evir)
    e2 = ctx.mbpt3_2(f.R, V, A, o, v)
    print("NT2", len(A), "E3_1", e1_ref, e1 - e1_ref, "E3_2", e2_ref, e2 - e2_ref)
    assert abs(e1 - e1_ref) <= 1e-10 * abs(e1_ref)
    assert abs(e2 - e2_ref) <= 1e-10 * abs(e2_ref)
    st = ctx.last_mbpt3_stage_ms()
    assert st["mbpt3"][1] == 1


def test_mbpt3_1_split_i8(ctx):
    o, v, n = 12, 60, 324
    f, x, w, A, V = _amps(o, v, n, 12, 31)
    e_ref = mbpt3.mbpt3_1(f.R, V, A, o, v, f.eocc, f.evir)
    e = ctx.mbpt3_1(V, A, f.R, f.eocc, f.evir, precision="split_i8")
    print("split-int8 MBPT3-1", e_ref, e - e_ref)
    assert abs(e - e_ref) <= 1e-7


def test_mbpt3_device_chain(ctx):
    """GPU amplitudes (brpa_t2_compute) kept as torch tensors in HBM, then both
    corrections; compared with the oracle chain on the same guess vectors."""
    import torch
    o, v, n, nf = 8, 50, 150, 12
    f = synth.make_fragment(o, v, n, 4)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=9)
    Rd = torch.from_numpy(f.R).cuda()
    A, V = ctx.t2(Rd, f.eocc, f.evir, x, w, seed=9, device=True)
    assert len(A) == len(A_ref)
    e1 = ctx.mbpt3_1(V, A, Rd, f.eocc, f.evir)
    e2 = ctx.mbpt3_2(Rd, V, A, o, v)
    e1_ref = mbpt3.mbpt3_1(f.R, V_ref, A_ref, o, v, f.eocc, f.evir)
    e2_ref = mbpt3.mbpt3_2(f.R, V_ref, A_ref, o, v)
    assert abs(e1 - e1_ref) <= 1e-8 * abs(e1_ref)
    assert abs(e2 - e2_ref) <= 1e-8 * abs(e2_ref)


def test_mbpt3_empty_amplitudes(ctx):
    f = synth.make_fragment(2, 3, 4, 1)
    V = np.zeros((6, 0))
    A = np.zeros(0)
    assert ctx.mbpt3_1(V, A, f.R, f.eocc, f.evir) == 0.0
    assert ctx.mbpt3_2(f.R, V, A, 2, 3) == 0.0


def test_mbpt3_1_indefinite_x_takes_the_eigen_route(ctx):
    """Amplitudes of both signs (orthonormal V, A in (-0.3, 0.3)) give an
    indefinite X, so -X has no Cholesky factor and the eigenpairs of X carry the
    contraction; the energy still matches the oracle."""
    o, v, n = 4, 20, 30
    f = synth.make_fragment(o, v, n, 12)
    rng = np.random.default_rng(5)
    V, _ = np.linalg.qr(rng.standard_normal((o * v, 25)))
    V = np.ascontiguousarray(V)
    A = rng.uniform(-0.3, 0.3, 25)
    Am, Bm = mbpt3.mbpt3_1_matrices(f.R, V, A, f.eocc, f.evir)
    X = mbpt3.lyapunov(Bm, Am)
    ev = np.linalg.eigvalsh(0.5 * (X + X.T))
    assert ev.min() < 0.0 < ev.max()
    e_ref = -mbpt3.trxk(f.R, V, X, o, v)
    e = ctx.mbpt3_1(V, A, f.R, f.eocc, f.evir)
    print("indefinite X: E", e_ref, e - e_ref)
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
