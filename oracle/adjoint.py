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


def ecrpa_grad_R_entries(R, eocc, evir, freqs, weights, entries, Ms=None):
    """dE/dR at selected (ai, P) entries only, for large inputs: needs the N x N
    inverses but only the selected rows of R G_k."""
    gaps = (evir[None, :] - eocc[:, None]).ravel()
    n = R.shape[1]
    Ms = inverses(R, eocc, evir, freqs) if Ms is None else Ms
    rows = sorted({ai for ai, _ in entries})
    out = {}
    acc = {ai: np.zeros(n) for ai in rows}
    for k, (x, w) in enumerate(zip(freqs, weights)):
        c = 4.0 * gaps / (gaps ** 2 + x ** 2)
        G = (w / (2.0 * math.pi)) * (Ms[k] - np.eye(n))
        for ai in rows:
            acc[ai] += 2.0 * c[ai] * (R[ai] @ G)
    for ai, p in entries:
        out[(ai, p)] = acc[ai][p]
    return out


def inverses(R, eocc, evir, freqs, Pis=None):
    """M_k = (1 + Pi_k)^{-1} for every frequency (Pi_k built here unless given)."""
    gaps = (evir[None, :] - eocc[:, None]).ravel()
    n = R.shape[1]
    Ms = []
    for k, x in enumerate(freqs):
        c = 4.0 * gaps / (gaps ** 2 + x ** 2)
        P = Pis[k] if Pis is not None else R.T @ (c[:, None] * R)
        Ms.append(np.linalg.inv(np.eye(n) + P))
    return Ms


def central_fd_R(R, eocc, evir, freqs, weights, entries, h=1e-5, Ms=None):
    """[E(R + h e_aiP) - E(R - h e_aiP)] / 2h, evaluated without cancellation.

    Changing one entry of row ai changes Pi_k by c_k(ai)(r' r'^T - r r^T), a
    rank-2 update, so with M_k = (1 + Pi_k)^{-1}, m = (M_k r)_P, a = r^T M_k r:
      det(1 + Pi_k') / det(1 + Pi_k) = 1 + delta,
      delta = c (+-2 h m + h^2 M_PP) + c^2 h^2 (m^2 - a M_PP)
    (matrix determinant lemma), and tr Pi_k' - tr Pi_k = c (+-2 h R_aiP + h^2).
    Ms: optional precomputed inverses (``inverses``)."""
    gaps = (evir[None, :] - eocc[:, None]).ravel()
    Ms = inverses(R, eocc, evir, freqs) if Ms is None else Ms
    out = {e: 0.0 for e in entries}
    for k, (x, w) in enumerate(zip(freqs, weights)):
        c = 4.0 * gaps / (gaps ** 2 + x ** 2)
        M = Ms[k]
        for (ai, p) in entries:
            r = R[ai]
            Mr = M @ r
            a = r @ Mr
            m = Mr[p]
            cc = c[ai]
            base = cc * cc * h * h * (m * m - a * M[p, p]) + cc * h * h * M[p, p]
            dp = cc * 2.0 * h * m + base
            dm = -cc * 2.0 * h * m + base
            dld = np.log1p(dp) - np.log1p(dm)
            dtr = cc * 4.0 * h * r[p]
            out[(ai, p)] += w / (2.0 * math.pi) * (dld - dtr)
    return {e: v / (2.0 * h) for e, v in out.items()}


def central_fd_eps(R, eocc, evir, freqs, weights, which, idx, h=1e-5, Ms=None):
    """[E(eps + h e) - E(eps - h e)] / 2h for one occupied (which='o') or virtual
    ('v') orbital energy, without cancellation: the change of c_k on the rows S
    of that orbital is a rank-|S| update, so
      log det(1 + Pi') - log det(1 + Pi) = log det(1 + diag(dc_S) R_S M R_S^T)."""
    nocc, nvir = len(eocc), len(evir)
    n = R.shape[1]
    rows = (np.arange(nvir) + idx * nvir) if which == "o" else (np.arange(nocc) * nvir + idx)
    gaps = (evir[None, :] - eocc[:, None]).ravel()
    Ms = inverses(R, eocc, evir, freqs) if Ms is None else Ms
    tot = 0.0
    for k, (x, w) in enumerate(zip(freqs, weights)):
        c = 4.0 * gaps / (gaps ** 2 + x ** 2)
        M = Ms[k]
        RS = R[rows]
        S = RS @ M @ RS.T
        d = gaps[rows]
        dd = (-h if which == "o" else h)
        terms = []
        for sgn in (1.0, -1.0):
            g2 = d + sgn * dd
            dc = 4.0 * g2 / (g2 ** 2 + x ** 2) - c[rows]
            X = np.eye(len(rows)) + dc[:, None] * S
            ld = np.linalg.slogdet(X)[1]
            dtr = np.sum(dc * np.sum(RS * RS, axis=1))
            terms.append(ld - dtr)
        tot += w / (2.0 * math.pi) * (terms[0] - terms[1])
    return tot / (2.0 * h)
