"""Direct-RPA correlation energy in the MO basis (oracle restatement).

* ``gpiug_mo``       -- rpa_GPiUG_MO, src/rpa/rpa_core_MO.f90:347-406:
                        DCRG(:,ai) = W_ai CRG(:,ai) with W_ai = d/(d^2+u^2)
                        (390-401), then GPiUG += Prefactor * CRG * DCRG^T
                        (dgemm 'N','C', 402-403); Prefactor = 4 closed shell,
                        2 per spin when NSpins > 1 (379-383).
* ``log_det_qr``     -- real_LogDet, src/linear-algebra/real_linalg.f90:793-825:
                        dgeqrf, then sum log|R_kk| in extended precision.
* ``ecrpa_analytic`` -- rpa_CC_EcRPA_Analytic, src/rpa/rpa_CC.f90:412-471:
                        per frequency, extended-precision trace, W = Pi + 1,
                        E(u) = logdet - tr, Ec += w_u E(u) / (2 pi).

The reference accumulates in REAL(F128); here numpy.longdouble (x87 80-bit
on x86-64) is used for the same sums.

Pinned by (tests/test_oracle_erpa.py): brute-force triple loop for Pi; QR vs
Cholesky log-det; the Sylvester determinant identity; the plasmon closed form
of the infinite-quadrature energy; the second-order (direct MP2) limit; the
lambda-integral identity of rpa_CC_VcLambdaPiU (src/rpa/rpa_CC.f90:341-361).
"""
import math

import numpy as np


def chi_weights(gaps, freq, prefactor=4.0):
    """Prefactor * d / (d^2 + u^2), the diagonal of the scaled GEMM."""
    return prefactor * gaps / (gaps ** 2 + freq ** 2)


def gpiug_mo(R, gaps, freq, prefactor=4.0, Pi=None):
    """Pi(u) (+)= Prefactor * R^T diag(d/(d^2+u^2)) R for R[ai][P]."""
    w = gaps / (gaps ** 2 + freq ** 2)
    dcrg = w[:, None] * R
    out = prefactor * (R.T @ dcrg)
    if Pi is None:
        return out
    Pi += out
    return Pi


def gpiug_mo_bruteforce(R, gaps, freq, prefactor=4.0):
    """Explicit loops; tiny inputs only (pins gpiug_mo)."""
    nai, n = R.shape
    out = np.zeros((n, n))
    for p in range(n):
        for q in range(n):
            s = 0.0
            for ai in range(nai):
                s += R[ai, p] * R[ai, q] * gaps[ai] / (gaps[ai] ** 2 + freq ** 2)
            out[p, q] = prefactor * s
    return out


def log_det_qr(W):
    """log|det W| from the R factor of a Householder QR, summed in long double."""
    r = np.linalg.qr(W, mode="r")
    d = np.abs(np.diag(r)).astype(np.longdouble)
    return np.sum(np.log(d))


def log_det_cholesky(W):
    """2 sum log L_kk of the Cholesky factor (valid for SPD W)."""
    L = np.linalg.cholesky(W)
    return 2.0 * np.sum(np.log(np.diag(L).astype(np.longdouble)))


def ecrpa_analytic(Pis, freqs, weights, logdet=log_det_qr, return_terms=False):
    """E_RPA = sum_u w_u/(2 pi) [log det(1+Pi_u) - tr Pi_u] from stored Pi(u)."""
    n = Pis[0].shape[0]
    e = np.longdouble(0.0)
    terms = []
    for u in range(len(freqs)):
        P = Pis[u]
        tr = np.sum(np.diag(P).astype(np.longdouble))
        W = P + np.eye(n)
        ld = logdet(W)
        eu = ld - tr
        terms.append(float(eu))
        e += np.longdouble(1.0 / (2.0 * math.pi)) * eu * np.longdouble(weights[u])
    if return_terms:
        return float(e), terms
    return float(e)


def ecrpa(R, gaps, freqs, weights, prefactor=4.0, logdet=log_det_qr, return_terms=False):
    """Fused oracle: build Pi(u) per frequency and accumulate E_RPA."""
    n = R.shape[1]
    e = np.longdouble(0.0)
    terms = []
    for u, (x, w) in enumerate(zip(freqs, weights)):
        P = gpiug_mo(R, gaps, x, prefactor)
        tr = np.sum(np.diag(P).astype(np.longdouble))
        P[np.diag_indices(n)] += 1.0
        eu = logdet(P) - tr
        terms.append(float(eu))
        e += np.longdouble(1.0 / (2.0 * math.pi)) * eu * np.longdouble(w)
    if return_terms:
        return float(e), terms
    return float(e)


def pack_lower(A):
    """LAPACK lower-packed storage (dtrttp, uplo 'L') of each N x N matrix in A
    (F, N, N) read column-major, i.e. A[w] holds Fortran A(i, j) at [j][i]
    (rpa_Pack / real_Pack, src/rpa/rpa.f90:379-387, linear-algebra/real_linalg.f90:737-750)."""
    A = np.asarray(A)
    F, n, _ = A.shape
    out = np.empty((F, n * (n + 1) // 2))
    for w in range(F):
        k = 0
        for j in range(n):
            out[w, k:k + n - j] = A[w, j, j:]      # column j, rows i >= j
            k += n - j
    return out


def unpack_lower(AP, n):
    """Full symmetric matrices of lower-packed AP (F, N(N+1)/2) (rpa_Unpack, rpa.f90:390-404)."""
    AP = np.asarray(AP)
    F = AP.shape[0]
    out = np.zeros((F, n, n))
    for w in range(F):
        k = 0
        for j in range(n):
            out[w, j, j:] = AP[w, k:k + n - j]
            out[w, j:, j] = AP[w, k:k + n - j]
            k += n - j
    return out
