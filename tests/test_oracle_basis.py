"""Pins of the rpa_G basis oracle (oracle/basis.py), CPU only."""
import numpy as np

from oracle import basis


def _spd(n, rank, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, rank)) * np.geomspace(1.0, 1e-6, rank)[None, :]
    return X @ X.T


def test_eigen_basis_is_the_spectral_decomposition():
    R = _spd(40, 40, 1)
    G, R2, nv, err = basis.rpa_g(R, basis.BASIS_EIGEN, 1e-3)
    lam = np.diag(R2)
    assert np.all(np.diff(lam) <= 1e-15)                       # largest first
    assert np.allclose(G.T @ G, np.eye(40), atol=1e-12)
    assert np.allclose((G * lam) @ G.T, R, atol=1e-12)
    assert err < 1e-3 and nv < 40


def test_nvecs_rule_is_the_shortest_prefix_with_small_tail():
    rng = np.random.default_rng(3)
    for _ in range(100):
        d = np.sort(10.0 ** rng.uniform(-9, 0, 30))[::-1]
        t = 10.0 ** rng.uniform(-8, 0)
        nv, err = basis.nvecs_rule(d, t)
        # dropped columns nv+1..N (1-based) and column nv itself are summed in err
        tails = [d[p:].sum() for p in range(30)]
        ok = [p for p in range(30) if tails[p] < t]
        want = (min(ok) + 1) if ok else 30
        assert nv == want


def test_random_basis_captures_a_low_rank_range():
    R = _spd(60, 8, 5)
    G, R2, nv, err = basis.rpa_g(R, basis.BASIS_RANDOM, 1e-8, seed=11)
    assert np.allclose(G.T @ G, np.eye(60), atol=1e-12)
    assert abs(np.trace(R2) - np.trace(R)) < 1e-12 * np.trace(R)
    assert np.max(np.abs(np.diag(R2)[8:])) < 1e-12 * np.trace(R)
    assert nv <= 8 + 1


def test_gaussian_generator_statistics():
    g = basis.gaussian(7, 300)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert np.array_equal(g, basis.gaussian(7, 300))
