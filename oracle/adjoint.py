"""Adjoint of the direct-RPA energy (oracle; no reference counterpart).

The reference computes no RPA gradients.  For
  E = sum_k w_k/(2 pi) [log det(1 + Pi_k) - tr Pi_k],
  Pi_k = R^T diag(c_k) R,  c_k = 4 d / (d^2 + u_k^2),  d_ai = e_a - e_i,
with the frequency grid held fixed (in the full program the grid's alpha
depends on the gaps through the histogram scan; that dependence is
piecewise constant and is not differentiated):
  G_k      = w_k/(2 pi) [(1 + Pi_k)^{-1} - 1]          (= dE/dPi_k, symmetric)
  dE/dR    = sum_k 2 diag(c_k) R G_k
  dE/dd_ai = sum_k (R G_k R^T)_{ai,ai} * 4 (u_k^2 - d^2)/(d^2 + u_k^2)^2
  dE/de_a  = sum_i dE/dd_ai,   dE/de_i = -sum_a dE/dd_ai.

Pinned by (tests/test_oracle_adjoint.py): central finite differences of
``oracle.erpa.ecrpa`` on random entries.
"""
import math

import numpy as np


def ecrpa_grad(R, eocc, evir, freqs, weights):
    nocc, nvir = len(eocc), len(evir)
    gaps = (evir[None, :] - eocc[:, None]).ravel()
    n = R.shape[1]
    dR = np.zeros_like(R)
    dd = np.zeros_like(gaps)
    for x, w in zip(freqs, weights):
        c = 4.0 * gaps / (gaps ** 2 + x ** 2)
        P = R.T @ (c[:, None] * R)
        G = (w / (2.0 * math.pi)) * (np.linalg.inv(np.eye(n) + P) - np.eye(n))
        RG = R @ G
        dR += 2.0 * c[:, None] * RG
        dd += np.sum(RG * R, axis=1) * 4.0 * (x ** 2 - gaps ** 2) / (gaps ** 2 + x ** 2) ** 2
    dd2 = dd.reshape(nocc, nvir)
    return dR, -dd2.sum(axis=1), dd2.sum(axis=0)
