"""Closed forms that pin the oracle independently of its own formulas.

* ``plasmon_energy`` -- infinite-quadrature direct-RPA energy
      E = 1/2 tr[(D^{1/2}(D + 4V)D^{1/2})^{1/2} - D - 2V],  V = R R^T (ov x ov),
  the direct-RPA specialisation (exchange integrals dropped) of the plasmon
  formula E = 1/2 sum(Omega - A_qq) in src/ccsd/arxiv/rpasimple.f90:89-143.
* ``dmp2_energy``    -- second-order limit of the same energy,
      E2 = -2 sum_{ia,jb} V_{ia,jb}^2 / (d_ia + d_jb).
* ``riccati_t2``     -- direct-ring amplitudes at exact frequency integration,
      T = 1/2 (1+K)^{-1}(1-K),  K = D^{-1/2} Omega D^{-1/2},
      Omega = (D^{1/2}(D + 4V)D^{1/2})^{1/2},
  the solution of B + A T' + T' A + T' B T' = 0 (A = D+2V, B = 2V) halved.
* ``ecrpa_lambda_integral`` -- E_RPA = int_0^1 VcLambda dl with
      VcLambda = -1/(2 pi) sum_u w_u sum_k l pi_k^2 / (1 + l pi_k),
  the reference's own cross-check (rpa_CC_VcLambdaPiU, src/rpa/rpa_CC.f90:341-361).
"""
import math

import numpy as np


def _sqrtm_spd(M):
    w, U = np.linalg.eigh(M)
    return (U * np.sqrt(np.maximum(w, 0.0))[None, :]) @ U.T


def plasmon_energy(R, gaps):
    V = R @ R.T
    sd = np.sqrt(gaps)
    M = sd[:, None] * (np.diag(gaps) + 4.0 * V) * sd[None, :]
    Om = _sqrtm_spd(M)
    return 0.5 * (np.trace(Om) - np.sum(gaps) - 2.0 * np.trace(V))


def dmp2_energy(R, gaps):
    V = R @ R.T
    return -2.0 * np.sum(V ** 2 / (gaps[:, None] + gaps[None, :]))


def riccati_t2(R, gaps):
    V = R @ R.T
    sd = np.sqrt(gaps)
    M = sd[:, None] * (np.diag(gaps) + 4.0 * V) * sd[None, :]
    Om = _sqrtm_spd(M)
    K = Om / sd[:, None] / sd[None, :]
    n = len(gaps)
    return 0.5 * np.linalg.solve(np.eye(n) + K, np.eye(n) - K)


def ecrpa_lambda_integral(Pis, weights, nlam=40):
    t, wl = np.polynomial.legendre.leggauss(nlam)
    lam = 0.5 * (t + 1.0)
    wl = 0.5 * wl
    eig = [np.linalg.eigvalsh(P) for P in Pis]
    e = 0.0
    for l, wlam in zip(lam, wl):
        vc = 0.0
        for u, w in enumerate(weights):
            p = eig[u]
            vc += w * np.sum(l * p ** 2 / (1.0 + l * p))
        e += wlam * (-vc / (2.0 * math.pi))
    return e
