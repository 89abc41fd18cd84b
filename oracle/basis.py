"""Basis of the effective dielectric matrix, rpa_G (oracle restatement; test
infrastructure, not the product path).

Reference: src/rpa/rpa.f90:31-173 (rpa_G); basis types
src/rpa/rpa_definitions.f90:92-104; QR = dgeqrf + dorgqr
(linear-algebra/real_linalg.f90:850-886).

  EIGEN (2):  G = eigenvectors of RWR, largest eigenvalue first; RWR <- diag
  RANDOM (1): G Gaussian, two subspace iterations G <- Q(RWR G); RWR <- G^T RWR G
  FULL (3):   G = I, NVecs = N
  NVecs: p = N..1, add diag(RWR)[p] while the sum stays below the target; NVecs
         is the last p added (1-based), N when none is.

The Gaussian numbers: Box-Muller on SplitMix64 counter draws (the same bits the
HIP library draws; the reference uses its own generator, so the random basis
itself cannot match the reference's).

Pinned by (tests/test_oracle_basis.py): EIGEN against the spectral
decomposition (G orthonormal, G diag(l) G^T = RWR); the NVecs rule against its
definition as the smallest prefix whose dropped tail sums below the target;
RANDOM on a rank-r RWR, where two subspace iterations capture the range
exactly (trace kept, trailing diagonal ~0); FULL as the identity.
"""
import math
from fractions import Fraction

import numpy as np

BASIS_RANDOM, BASIS_EIGEN, BASIS_FULL_CHOLESKY = 1, 2, 3
_M64 = (1 << 64) - 1


def _splitmix(seed, ctr):
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + ctr.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def gaussian(seed, n):
    """n x n standard normal matrix, column-major element e = i + n j."""
    e = np.arange(n * n, dtype=np.uint64)
    u1 = ((_splitmix(seed, 2 * e + 1) >> np.uint64(11)).astype(np.float64) + 1.0) / 2.0 ** 53
    u2 = (_splitmix(seed, 2 * e + 2) >> np.uint64(11)).astype(np.float64) / 2.0 ** 53
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    return g.reshape((n, n), order="F")


def nvecs_rule(diag, target):
    """(NVecs, AbsError) of rpa.f90:141-152 with exact sums."""
    n = len(diag)
    nv, acc = n, Fraction(0)
    for p in range(n - 1, -1, -1):
        nxt = acc + Fraction(float(diag[p]))
        if nxt < Fraction(float(target)):
            acc = nxt
            nv = p + 1
        else:
            break
    return nv, float(acc)


def rpa_g(RWR, basis, target, seed=0):
    """(G, RWR', NVecs, AbsError) with column-major math arrays G[i, j]."""
    n = RWR.shape[0]
    if basis == BASIS_FULL_CHOLESKY:
        return np.eye(n), RWR.copy(), n, 0.0
    if basis == BASIS_EIGEN:
        lam, G = np.linalg.eigh(-RWR)
        lam = -lam
        R2 = np.diag(lam)
    elif basis == BASIS_RANDOM:
        G = gaussian(seed, n)
        for _ in range(2):
            G, _r = np.linalg.qr(RWR @ G)
        R2 = G.T @ (RWR @ G)
    else:
        raise ValueError("invalid basis type")
    nv, err = nvecs_rule(np.diag(R2), target)
    return G, R2, nv, err
