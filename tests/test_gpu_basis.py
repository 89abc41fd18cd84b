"""GPU parity of the rpa_G basis (brpa_rpa_g) against the CPU oracle
(oracle/basis.py): NVecs equal; eigenvalues 1e-12 relative; eigenvectors up to
sign; the random basis on the same Gaussian draws within 1e-9."""
import numpy as np
import pytest

from brpa import synth
from oracle import basis, erpa, grid

pytestmark = pytest.mark.gpu


def _rwr(o, v, n, seed):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 8)
    return erpa.gpiug_mo(f.R, f.gaps(), x[0])


@pytest.mark.parametrize("o,v,n", [(2, 3, 4), (6, 40, 90), (8, 72, 400)])
def test_rpa_g_eigen(ctx, o, v, n):
    R = _rwr(o, v, n, 5)
    target = 1e-3 * np.trace(R)
    G_ref, R2_ref, nv_ref, err_ref = basis.rpa_g(R, basis.BASIS_EIGEN, target)
    G, R2, nv, err = ctx.rpa_g(R, "eigen", target)
    Gm = G.T                                        # [j][i] -> math (i, j)
    lam, lam_ref = np.diag(R2), np.diag(R2_ref)
    print("NVecs", nv, nv_ref, "of", n)
    assert nv == nv_ref
    assert np.max(np.abs(lam - lam_ref)) <= 1e-12 * np.max(np.abs(lam_ref))
    assert abs(err - err_ref) <= 1e-12 * max(abs(err_ref), 1e-300) + 1e-15
    # columns of well-separated eigenvalues match up to sign
    gap = np.abs(np.diff(lam_ref))
    sep = [k for k in range(min(nv, 20)) if (k == 0 or gap[k - 1] > 1e-6 * lam_ref[0])
           and (k == n - 1 or gap[k] > 1e-6 * lam_ref[0])]
    for k in sep:
        s = np.sign(Gm[:, k] @ G_ref[:, k])
        assert np.max(np.abs(s * Gm[:, k] - G_ref[:, k])) < 1e-8
    assert np.allclose(Gm.T @ Gm, np.eye(n), atol=1e-12)


def test_rpa_g_random_and_full(ctx):
    R = _rwr(6, 40, 90, 7)
    target = 1e-6 * np.trace(R)
    G_ref, R2_ref, nv_ref, _ = basis.rpa_g(R, basis.BASIS_RANDOM, target, seed=3)
    G, R2, nv, _ = ctx.rpa_g(R, "random", target, seed=3)
    assert nv == nv_ref
    assert np.max(np.abs(G.T - G_ref)) < 1e-9
    assert np.max(np.abs(R2.T - R2_ref)) < 1e-9 * np.max(np.abs(R2_ref))
    G, R2, nv, err = ctx.rpa_g(R, "full", target)
    assert nv == 90 and err == 0.0 and np.array_equal(G, np.eye(90)) and np.array_equal(R2, R)
