"""Third-order exchange corrections MBPT3-1 and MBPT3-2 from factorised
amplitudes (oracle restatement; test infrastructure, not the product path).

Reference: src/rpa/rpa_CC_Exchange.f90
  rpa_CC_Ukl 178-207, rpa_CC_Dkl 210-229, rpa_CC_Exchange_MBPT3_1 232-270,
  rpa_CC_TrXK 273-336, rpa_CC_TrW2 339-378, rpa_CC_Exchange_MBPT3_2 381-393;
  real_Lyapunov src/linear-algebra/real_linalg.f90:125-171 (Schur + dtrsyl).
Callers: src/rpa/rpa_CC.f90:977-997 (MBPT3_1 alone, or MBPT3_1 + MBPT3_2).

MBPT3-1 (amplitudes T = V diag(A) V^T, V[ai][mu] with orthonormal columns):
  U = (R^T V)^T (R^T V)            U_kl = sum V_k,ai (ai|bj) V_l,bj
  D = V^T diag(d) V                d_ai = e_a - e_i
  Amat_mn = (A_m + A_n) D_mn,  Bmat_mn = D_mn + 2 U_mn + 4 U_mn A_n
  solve Bmat^T X + X Bmat = Amat   (X = V^T (dT/dlambda) V)
  E = -TrXK,  TrXK = sum_{aibj} (aj|bi) (V X V^T)(ai, bj)
MBPT3-2:
  E = 2 sum_{c,k,l} ( sum_{a,i} R(l, ai) T(ak, ci) )^2

Pinned by (tests/test_oracle_mbpt3.py): the Lyapunov solution against a
Kronecker-product solve on small matrices; TrXK against explicit ov x ov
tensors; MBPT3-1 against a central finite difference in the coupling
strength, d/dlambda of the SOSEX energy of T(lambda) at lambda = 1 (the
reference's own numerical variant, rpa.f90:1589-1604); MBPT3-2 against the
explicit contraction with the ov x ov amplitude tensor.
"""
import numpy as np
import scipy.linalg

from . import sosex as osx


def mbpt3_1_matrices(R, V, A, eocc, evir):
    """(Amat, Bmat) of rpa_CC_Exchange_MBPT3_1 (rpa_CC_Exchange.f90:248-262)."""
    W = R.T @ V                                   # (N, NT2)
    U = W.T @ W
    d = (np.asarray(evir)[None, :] - np.asarray(eocc)[:, None]).ravel()
    D = V.T @ (d[:, None] * V)
    Amat = (A[:, None] + A[None, :]) * D
    Bmat = D + 2.0 * U + 4.0 * U * A[None, :]
    return Amat, Bmat


def lyapunov(Bmat, Amat):
    """X with Bmat^T X + X Bmat = Amat (Bartels-Stewart, as real_Lyapunov)."""
    return scipy.linalg.solve_continuous_lyapunov(Bmat.T, Amat)


def trxk(R, V, X, nocc, nvir):
    """rpa_CC_TrXK (273-336): pair loop over j <= i."""
    naux, nt = R.shape[1], V.shape[1]
    R3 = R.reshape(nocc, nvir, naux)
    V3 = V.reshape(nocc, nvir, nt)
    t = 0.0
    for j in range(nocc):
        for i in range(j, nocc):
            Wij = R3[i].T @ V3[j]
            if i != j:
                Wji = R3[j].T @ V3[i]
                t += 2.0 * np.sum(X * (Wji.T @ Wij))
            else:
                t += np.sum(X * (Wij.T @ Wij))
    return t


def mbpt3_1(R, V, A, nocc, nvir, eocc, evir):
    Amat, Bmat = mbpt3_1_matrices(R, V, A, eocc, evir)
    X = lyapunov(Bmat, Amat)
    return -trxk(R, V, X, nocc, nvir)


def mbpt3_2(R, V, A, nocc, nvir):
    """2 TrW2 following rpa_CC_TrW2 (339-378): loops over c and k."""
    naux, nt = R.shape[1], V.shape[1]
    V3 = V.reshape(nocc, nvir, nt)                # [i][a][mu]
    tr = 0.0
    for c in range(nvir):
        Vli = A[:, None] * V3[:, c, :].T          # (NT2, o): A_mu V(mu, c, i)
        for k in range(nocc):
            Wai = V3[k] @ Vli                     # (v, o): sum_mu V(mu,a,k) A_mu V(mu,c,i)
            Wm = R.T @ Wai.T.ravel()              # ai = i*v + a
            tr += np.sum(Wm * Wm)
    return 2.0 * tr


def trxk_bruteforce(R, V, X, nocc, nvir):
    """sum_{aibj} (aj|bi) (V X V^T)(ai, bj) with explicit tensors (tiny inputs)."""
    nai = nocc * nvir
    G = R @ R.T                                   # (ai|bj)
    T = V @ X @ V.T
    t = 0.0
    for i in range(nocc):
        for a in range(nvir):
            for j in range(nocc):
                for b in range(nvir):
                    t += G[j * nvir + a, i * nvir + b] * T[i * nvir + a, j * nvir + b]
    return t


def mbpt3_2_bruteforce(R, V, A, nocc, nvir):
    """2 sum_{c,k,l} (sum_{a,i} R(l, ai) T(ak, ci))^2 with the explicit T (tiny inputs)."""
    T = (V * A[None, :]) @ V.T
    naux = R.shape[1]
    tr = 0.0
    for c in range(nvir):
        for k in range(nocc):
            for l in range(naux):
                s = 0.0
                for a in range(nvir):
                    for i in range(nocc):
                        s += R[i * nvir + a, l] * T[k * nvir + a, i * nvir + c]
                tr += s * s
    return 2.0 * tr


def lyapunov_sign(Bmat, Amat, tol=1e-14, maxit=100):
    """The same X by the scaled matrix-sign iteration the HIP library uses
    (B <- (cB + (cB)^-1)/2, A <- (cA + c^-1 B^-T A B^-1)/2, X = A_inf / 2);
    kept here to check the iteration itself against Bartels-Stewart."""
    B, Q = Bmat.copy(), Amat.copy()
    n = B.shape[0]
    for it in range(maxit):
        Bi = np.linalg.inv(B)
        c = np.sqrt(np.linalg.norm(Bi) / np.linalg.norm(B))
        Qn = 0.5 * (c * Q + (Bi.T @ Q @ Bi) / c)
        Bn = 0.5 * (c * B + Bi / c)
        done = np.linalg.norm(Bn - np.eye(n)) <= tol * np.sqrt(n) * 10
        B, Q = Bn, Qn
        if done:
            break
    return 0.5 * Q, it + 1
