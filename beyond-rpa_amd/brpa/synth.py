"""Seeded synthetic fragments for the correlation-energy workloads.

A fragment is what the reference's MO-basis RPA path holds after the AO->MO
Cholesky transform (src/rpa/rpa_core_MO.f90:220-301) and the core freeze /
Fermi shift (src/rpa/rpa.f90:1834-1851):

* ``R``    -- Cholesky vectors, C row-major ``R[ai][P]`` of shape (o*v, N), with
              ai = i*v + a (a fastest).  This is bit-identical to the Fortran
              column-major ``Rkai(P, ai)`` / ``CRG(P, ai)`` of the reference.
* ``eocc`` -- active occupied energies, ascending (Hartree).
* ``evir`` -- virtual energies, ascending (Hartree).

Distributions (fixed by the workload definitions): R[ai,P] = N(0,1) * 0.3 *
exp(-4 P / N) with P 0-based, eps_occ ~ U[-1.5, -0.3] sorted ascending,
eps_vir = geomspace(0.05, 3.0, v).  Draw order from one PCG64 stream: eps_occ
first, then R.
"""
from dataclasses import dataclass

import numpy as np

BASE_SEED = 20261202


@dataclass
class Fragment:
    nocc: int
    nvir: int
    naux: int
    eocc: np.ndarray
    evir: np.ndarray
    R: np.ndarray
    seed: int

    @property
    def nai(self):
        return self.nocc * self.nvir

    def gaps(self):
        """Delta_ai = eps_a - eps_i in the ai = i*v + a order."""
        return (self.evir[None, :] - self.eocc[:, None]).ravel()


def make_fragment(nocc, nvir, naux, seed=BASE_SEED):
    rng = np.random.Generator(np.random.PCG64(seed))
    eocc = np.sort(rng.uniform(-1.5, -0.3, nocc))
    evir = np.geomspace(0.05, 3.0, nvir)
    scale = 0.3 * np.exp(-4.0 * np.arange(naux, dtype=np.float64) / naux)
    R = rng.standard_normal((nocc * nvir, naux)) * scale[None, :]
    return Fragment(nocc, nvir, naux, eocc, evir, np.ascontiguousarray(R), seed)


def batch_shapes(nfrag=256, seed=BASE_SEED):
    """Shapes of the batched workload: o ~ U{8..40}, v ~ U{70..350}, N = round(4.5 (o+v)).

    The shapes come from one PCG64 stream seeded with ``seed``; fragment f's
    own contents use seed + 1 + f.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.integers(8, 41, nfrag)
    v = rng.integers(70, 351, nfrag)
    n = np.rint(4.5 * (o + v)).astype(np.int64)
    return [(int(a), int(b), int(c)) for a, b, c in zip(o, v, n)]


def make_batch(nfrag=256, seed=BASE_SEED, shapes=None):
    shapes = batch_shapes(nfrag, seed) if shapes is None else shapes
    return [make_fragment(o, v, n, seed + 1 + f) for f, (o, v, n) in enumerate(shapes)]


# Named workloads (nocc, nvir, naux, nfreqs)
CONFIGS = {
    "cpu_small": (8, 72, 400, 16),
    "single_large": (36, 300, 1800, 24),
    "adjoint": (48, 400, 2400, 16),
    "sosex_large": (64, 600, 3200, 16),
}
BATCH_NFREQS = 16


@dataclass
class ThcFragment:
    """Tensor-hypercontraction factors of one fragment (the THC route's inputs).

    ``Xga`` (nvir, ngrid), ``Xgi`` (nocc, ngrid), ``Zgk`` (naux, ngrid) are C
    arrays holding X(g, a) at [a][g]: the bytes of the reference's Fortran
    Xga(NGridTHC, NVirt), Xgi(NGridTHC, NOcc), Zgk(NGridTHC, NCholesky)
    (src/rpa/rpa_THC.f90:373-403).  Distributions: X ~ N(0, 1), Z(g, P) =
    N(0, 1) * 0.3 * exp(-4 P / N) / sqrt(ngrid), so that R(P, ai) = sum_g
    Z X X has the scale profile of the MO-route fragments.  Draw order from one
    PCG64 stream: eps_occ, Xgi, Xga, Zgk.
    """
    nocc: int
    nvir: int
    naux: int
    ngrid: int
    eocc: np.ndarray
    evir: np.ndarray
    Xga: np.ndarray
    Xgi: np.ndarray
    Zgk: np.ndarray
    seed: int

    def gaps(self):
        return (self.evir[None, :] - self.eocc[:, None]).ravel()


def make_thc_fragment(nocc, nvir, naux, ngrid, seed=BASE_SEED):
    rng = np.random.Generator(np.random.PCG64(seed))
    eocc = np.sort(rng.uniform(-1.5, -0.3, nocc))
    evir = np.geomspace(0.05, 3.0, nvir)
    Xgi = rng.standard_normal((nocc, ngrid))
    Xga = rng.standard_normal((nvir, ngrid))
    scale = 0.3 * np.exp(-4.0 * np.arange(naux, dtype=np.float64) / naux) / np.sqrt(ngrid)
    Zgk = rng.standard_normal((naux, ngrid)) * scale[:, None]
    return ThcFragment(nocc, nvir, naux, ngrid, eocc, evir, np.ascontiguousarray(Xga),
                       np.ascontiguousarray(Xgi), np.ascontiguousarray(Zgk), seed)


# THC-route workload (nocc, nvir, naux, ngrid, nfreqs): the single-fragment
# shape with a THC grid of 12 points per orbital
THC_CONFIG = (36, 300, 1800, 4032, 16)
