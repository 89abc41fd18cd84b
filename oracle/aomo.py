"""AO -> MO transform of the Cholesky vectors (oracle restatement; test
infrastructure, not the product path).

Reference: src/rpa/rpa_core_MO.f90:40-82 (rpa_Cab), 85-134 (rpa_Caa),
137-190 (rpa_C), 193-202 (rpa_CRGxy), 233-301 (rpa_CRG):

  CRG(P, ai) = sum_pq RG(P, pq) C(ai, pq)
  C(ai, pq)  = Co(p, i) Cv(q, a) + Co(q, i) Cv(p, a)   p != q
             = Co(p, i) Cv(p, a)                        p == q

Packed AO-pair columns, shell pair by shell pair: for shells A != B the
block column p + Na q (p over A, q over B, p fastest); for A == B the lower
triangle column by column (q outer, the diagonal (q, q) first, then p > q).

Arrays (math-shaped): RG[pq, P] (npq, nvecs), Co[p, i] (nao, nocc),
Cv[q, a] (nao, nvir); CRG is C R[ai][P] (ov, nvecs), ai = i*nvir + a.

Pinned by (tests/test_oracle_aomo.py): the packed two-term formula equals the
unpacked contraction sum_{p,q} RGfull(P, p, q) Co(p, i) Cv(q, a) over all
ordered AO pairs with RGfull symmetric in (p, q), the meaning of a Cholesky
basis of the symmetric AO-pair space; the pair table covers every unordered
AO pair exactly once.
"""
import numpy as np


def pair_table(shell_pairs, shell_loc, shell_size):
    """(npq, 2) int32 AO index pairs in the order of rpa_C / rpa_Cab / rpa_Caa.

    shell_pairs: list of (A, B) shell indices (0-based), shell_loc: first AO of
    each shell, shell_size: functions per shell.
    """
    out = []
    for A, B in shell_pairs:
        la, lb, na, nb = shell_loc[A], shell_loc[B], shell_size[A], shell_size[B]
        if A != B:
            for q in range(nb):
                for p in range(na):
                    out.append((la + p, lb + q))
        else:
            for q in range(na):
                out.append((la + q, la + q))
                for p in range(q + 1, na):
                    out.append((la + p, la + q))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def c_matrix(pairs, Co, Cv):
    """C(ai, pq) of rpa_Cab / rpa_Caa, (ov, npq)."""
    p, q = pairs[:, 0], pairs[:, 1]
    nocc, nvir = Co.shape[1], Cv.shape[1]
    # [i, a, pq]
    C = Co[p, :].T[:, None, :] * Cv[q, :].T[None, :, :]
    off = p != q
    C[:, :, off] += Co[q[off], :].T[:, None, :] * Cv[p[off], :].T[None, :, :]
    return C.reshape(nocc * nvir, len(p))


def crg(RG, pairs, Co, Cv):
    """R[ai][P] = sum_pq C(ai, pq) RG[pq, P]."""
    return c_matrix(pairs, Co, Cv) @ RG


def crg_unpacked(RG, pairs, Co, Cv):
    """sum over all ordered (p, q) of RGfull[p, q, P] Co(p, i) Cv(q, a), RGfull symmetric."""
    nao = Co.shape[0]
    full = np.zeros((nao, nao, RG.shape[1]))
    for k, (p, q) in enumerate(pairs):
        full[p, q] = RG[k]
        full[q, p] = RG[k]
    X = np.einsum("pqP,pi,qa->iaP", full, Co, Cv)
    return X.reshape(Co.shape[1] * Cv.shape[1], RG.shape[1])
