"""Exchange-type (SOSEX) correction from factorised amplitudes (oracle restatement).

Reference: rpa_CC_Exchange_SOSEX and rpa_CC_EcExchange1,
src/rpa/rpa_CC_Exchange.f90:36-116:

  EcExchange = -Alpha sum_{j<=i} (2 - delta_ij) sum_{P,mu} W_ji(P,mu) W_ij(P,mu) B_mu
  W_ij(P, mu) = sum_a R(P, a, i) V(mu, a, j)
  B = A, or A / (1 - 4 A^2) when S2Operator (the cumulant route).

which equals -Alpha sum_{aibj} (aj|bi) T(ai, bj) with T = V^T diag(B) V.

Layouts (C row-major, bit-identical to the reference's Fortran arrays):
R[ai][P] (Rkai(P, a, i)), V[ai][mu] (Vkai(mu, a, i)), ai = i*v + a.

Pinned by (tests/test_oracle_sosex.py): brute force over explicit ov x ov
tensors (aj|bi) and T(ai,bj) on tiny inputs.
"""
import numpy as np


def sosex(R, V, A, nocc, nvir, alpha=1.0, s2_operator=False):
    naux = R.shape[1]
    nt = V.shape[1]
    B = A / (1.0 - 4.0 * A ** 2) if s2_operator else A
    R3 = R.reshape(nocc, nvir, naux)       # [i][a][P]
    V3 = V.reshape(nocc, nvir, nt)         # [j][a][mu]
    e = 0.0
    for j in range(nocc):
        for i in range(j, nocc):
            Wij = R3[i].T @ V3[j]          # (N, NT2): sum_a R(P,a,i) V(mu,a,j)
            if i != j:
                Wji = R3[j].T @ V3[i]
                e += 2.0 * (-np.sum(Wji * Wij * B[None, :]))
            else:
                e += -np.sum(Wij * Wij * B[None, :])
    return alpha * e


def sosex_bruteforce(R, V, A, nocc, nvir, alpha=1.0):
    """-alpha sum_{aibj} (aj|bi) T(ai,bj) with explicit tensors (tiny inputs)."""
    T = (V * A[None, :]) @ V.T             # T[ai][bj]
    e = 0.0
    for i in range(nocc):
        for a in range(nvir):
            for j in range(nocc):
                for b in range(nvir):
                    ajbi = float(R[j * nvir + a] @ R[i * nvir + b])
                    e += ajbi * T[i * nvir + a, j * nvir + b]
    return -alpha * e
