"""GPU adjoint dE/dR, dE/d(eps) against the oracle's closed form and against
central finite differences (h = 1e-5).

E_RPA is 1e3-1e5 Ha here, so subtracting two FP64 energies that differ by
2h|g| loses ~1e-5 relative; the central differences are instead evaluated
without cancellation (oracle.adjoint.central_fd_*: matrix determinant lemma
for the rank-2 / rank-|S| change of Pi, exact trace difference).  They are
still [E(x+h) - E(x-h)] / 2h."""
import numpy as np
import pytest

from brpa import synth
from oracle import adjoint, erpa, grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("o,v,n", [(4, 20, 60), (5, 30, 150), (3, 50, 200)])
def test_grad_matches_oracle_closed_form(ctx, o, v, n):
    """One Cholesky block (N < 64) and several with a ragged last block."""
    f = synth.make_fragment(o, v, n, seed=21)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    dR, deo, dev = ctx.ecrpa_grad(f.R, f.eocc, f.evir, x, w)
    rR, ro, rv = adjoint.ecrpa_grad(f.R, f.eocc, f.evir, x, w)
    assert np.max(np.abs(dR - rR)) <= 1e-10 * np.max(np.abs(rR))
    assert np.max(np.abs(deo - ro)) <= 1e-10 * np.max(np.abs(ro))
    assert np.max(np.abs(dev - rv)) <= 1e-10 * np.max(np.abs(rv))


def _fd_worst(f, x, w, dR, deo, dev, ents, Ms, occ_idx, vir_idx):
    fd = adjoint.central_fd_R(f.R, f.eocc, f.evir, x, w, ents, 1e-5, Ms=Ms)
    worst = max(abs(fd[e] - dR[e]) / abs(fd[e]) for e in ents)
    for k in occ_idx:
        v = adjoint.central_fd_eps(f.R, f.eocc, f.evir, x, w, "o", k, 1e-5, Ms=Ms)
        worst = max(worst, abs(v - deo[k]) / abs(v))
    for k in vir_idx:
        v = adjoint.central_fd_eps(f.R, f.eocc, f.evir, x, w, "v", k, 1e-5, Ms=Ms)
        worst = max(worst, abs(v - dev[k]) / abs(v))
    return worst


def test_grad_central_fd_small(ctx):
    o, v, n = 6, 40, 150
    f = synth.make_fragment(o, v, n, seed=3)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 12)
    dR, deo, dev = ctx.ecrpa_grad(f.R, f.eocc, f.evir, x, w)
    rng = np.random.default_rng(3)
    ents = [(int(rng.integers(0, o * v)), int(rng.integers(0, n))) for _ in range(24)]
    Ms = adjoint.inverses(f.R, f.eocc, f.evir, x)
    worst = _fd_worst(f, x, w, dR, deo, dev, ents, Ms, range(3), (0, v // 2, v - 1))
    print("worst relative FD error", worst)
    assert worst <= 1e-6


def test_grad_adjoint_workload(ctx):
    """o=48, v=400, N=2400, 16 points: 64 random R entries and 6 orbital
    energies by central differences (h=1e-5), 8 R entries against the FP64
    closed form.  The N x N Pi(u) the CPU side inverts are built by the GPU
    (their parity is test_gpiug_mo_matches_oracle)."""
    torch = pytest.importorskip("torch")
    o, v, n, nf = 48, 400, 2400, 16
    f = synth.make_fragment(o, v, n, seed=20261202)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Rd = torch.from_numpy(f.R).cuda()
    dR, deo, dev = ctx.ecrpa_grad(Rd, f.eocc, f.evir, x, w)
    dR = dR.cpu().numpy()
    Pis = []
    for xx in x:
        P = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        ctx.gpiug_mo(P, Rd, f.eocc, f.evir, xx, 4.0)
        Pis.append(P.cpu().numpy())
    Ms = adjoint.inverses(f.R, f.eocc, f.evir, x, Pis=Pis)
    rng = np.random.default_rng(2024)
    ents = [(int(rng.integers(0, o * v)), int(rng.integers(0, n))) for _ in range(64)]
    worst = _fd_worst(f, x, w, dR, deo, dev, ents, Ms, (0, 23, 47), (0, 199, 399))
    print("worst relative FD error (adjoint workload)", worst)
    assert worst <= 1e-6
    ref = adjoint.ecrpa_grad_R_entries(f.R, f.eocc, f.evir, x, w, ents[:8], Ms=Ms)
    for e, g in ref.items():
        assert abs(dR[e] - g) <= 1e-9 * abs(g)


def test_autograd_function(ctx):
    torch = pytest.importorskip("torch")
    from brpa import autograd as bag
    f = synth.make_fragment(5, 30, 80, seed=8)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 10)
    R = torch.from_numpy(f.R.copy()).cuda().requires_grad_(True)
    eo = torch.from_numpy(f.eocc.copy()).requires_grad_(True)
    ev = torch.from_numpy(f.evir.copy()).requires_grad_(True)
    E = bag.ecrpa(R, eo, ev, x, w, context=ctx)
    assert abs(float(E.detach()) - erpa.ecrpa(f.R, f.gaps(), x, w)) <= 1e-8
    E.backward()
    rR, ro, rv = adjoint.ecrpa_grad(f.R, f.eocc, f.evir, x, w)
    assert np.allclose(R.grad.cpu().numpy(), rR, rtol=1e-9, atol=1e-12 * np.abs(rR).max())
    assert np.allclose(eo.grad.numpy(), ro, rtol=1e-9)
    assert np.allclose(ev.grad.numpy(), rv, rtol=1e-9)
