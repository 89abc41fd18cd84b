"""THC route: Cholesky vectors and Pi(u) from the tensor-hypercontraction
factors (oracle restatement; test infrastructure, not the product path).

Reference: src/rpa/rpa_THC.f90:215-282 (rpa_THC_PiU), 285-339
(rpa_THC_PiU_BlockAlgo), 342-370 (rpa_THC_ChiRkai), 373-403 (rpa_THC_Rkab),
405-422 (rpa_THC_Rkai, the per-occupied association).

  R(P, ai) = sum_g Zgk(g, P) Xga(g, a) Xgi(g, i),   ai = a + NVirt (i - 1)
  Pi(u)    = Alpha * sum_ai R(:, ai) d_ai / (d_ai^2 + u^2) R(:, ai)^T
  Alpha    = 4 for one spin block (closed shell), 2 per block otherwise

Arrays are the math-shaped transposes of the reference's: Xga[g, a],
Xgi[g, i], Zgk[g, P] (numpy (ngrid, n) arrays); R is C R[ai][P] (ov, naux),
the bytes of the reference's Rkai(P, ai).

Pinned by (tests/test_oracle_thc.py): a triple loop over (g, ai, P) on tiny
inputs; the per-occupied association of rpa_THC_Rkai (a different summation
order over the same index map); the MO-route Pi (gpiug_mo, itself pinned by
brute force and the Sylvester identity) of the produced R; Pi symmetric
positive semidefinite; the closed-shell prefactor equals two open-shell spin
blocks of the same orbitals.
"""
import numpy as np

from . import erpa


def thc_rkai(Xga, Xgi, Zgk):
    """R[i*v + a][P] = sum_g Xgi[g, i] Xga[g, a] Zgk[g, P] (rpa_THC_Rkab 373-403)."""
    G, v = Xga.shape
    o = Xgi.shape[1]
    H = (Xgi[:, :, None] * Xga[:, None, :]).reshape(G, o * v)   # Xgab(g, ai), a fastest
    return H.T @ Zgk


def thc_rkai_bruteforce(Xga, Xgi, Zgk):
    """Explicit loops; tiny inputs only (pins thc_rkai)."""
    G, v = Xga.shape
    o, N = Xgi.shape[1], Zgk.shape[1]
    R = np.zeros((o * v, N))
    for i in range(o):
        for a in range(v):
            for P in range(N):
                s = 0.0
                for g in range(G):
                    s += Zgk[g, P] * Xga[g, a] * Xgi[g, i]
                R[i * v + a, P] = s
    return R


def thc_rkai_per_occupied(Xga, Xgi, Zgk):
    """rpa_THC_Rkai (405-422): for each i, R_i = Zgk^T (Xga o Xgi(:, i))."""
    G, v = Xga.shape
    o, N = Xgi.shape[1], Zgk.shape[1]
    R = np.empty((o * v, N))
    for i in range(o):
        R[i * v:(i + 1) * v, :] = (Xga * Xgi[:, i:i + 1]).T @ Zgk
    return R


def thc_piu(Xga, Xgi, Zgk, eocc, evir, freqs, prefactor=4.0):
    """[Pi(u_w)] of one spin block (rpa_THC_PiU_BlockAlgo 285-339), plus R."""
    R = thc_rkai(Xga, Xgi, Zgk)
    gaps = (np.asarray(evir)[None, :] - np.asarray(eocc)[:, None]).ravel()
    return np.stack([erpa.gpiug_mo(R, gaps, u, prefactor) for u in freqs]), R
