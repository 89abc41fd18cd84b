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

Also rpa_CCD_corrections_1b2g (src/rpa/rpa_CCD_Corrections.f90:79-290),
thc_1b2g below.

Pinned by (tests/test_oracle_thc.py): a triple loop over (g, ai, P) on tiny
inputs; the per-occupied association of rpa_THC_Rkai (a different summation
order over the same index map); the MO-route Pi (gpiug_mo, itself pinned by
brute force and the Sylvester identity) of the produced R; Pi symmetric
positive semidefinite; the closed-shell prefactor equals two open-shell spin
blocks of the same orbitals.  thc_1b2g: Ec1b / 2 equals the SOSEX energy of
the same amplitudes with R built from the THC factors and Zgh = Zgk Zgk^T
(the reference's own statement, rpa_CCD_Corrections.f90:68-71); Ec2g against
explicit (ab|ij) tensors; the full-sum form against the reference's block walk.
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


def thc_1b2g(Zgh, Xgi, Yga, U, A):
    """rpa_CCD_corrections_1b2g (rpa_CCD_Corrections.f90:79-290).

    Zgh (G, G) symmetric, Xgi (G, o), Yga (G, v), U (v, o, NT2) math arrays
    (U[a, i, mu] = Uaim(a, i, mu)), A (NT2,).  Returns (Ec1b, Ec2g, EigSOSEX,
    Eig2g, YXUggm) with Ec1b before the caller's factor 1/2.  The reference's
    walk over lower-triangular grid blocks (diagonal blocks once, off-diagonal
    blocks with both orders) is the full sum over (g, h) for a symmetric Z.
    """
    G = Zgh.shape[0]
    nt = len(A)
    S2 = np.empty(nt)
    S1 = np.empty(nt)
    diag = np.empty((G, nt))
    for mu in range(nt):
        M = (Yga @ U[:, :, mu]) @ Xgi.T          # [YU] then [YXU](g, h)
        S2[mu] = np.sum(Zgh * M * M)
        S1[mu] = np.sum(Zgh * M * M.T)
        diag[:, mu] = np.diag(M)
    ec2g = 2.0 * np.sum(-2.0 * S2 * A ** 2)
    ec1b = 2.0 * np.sum(-S1 * A)
    return ec1b, ec2g, -S1 * A, -4.0 * S2 * A ** 2, diag


def thc_1b2g_blocked(Zgh, Xgi, Yga, U, A, block):
    """The reference's lower-triangular block walk itself (pins thc_1b2g's full sums)."""
    G = Zgh.shape[0]
    nb = -(-G // block)
    S2 = np.zeros(len(A))
    S1 = np.zeros(len(A))
    for mu in range(len(A)):
        M = (Yga @ U[:, :, mu]) @ Xgi.T
        for p in range(nb):
            for q in range(p + 1):
                g0, g1 = p * block, min((p + 1) * block, G)
                h0, h1 = q * block, min((q + 1) * block, G)
                W0, W1 = Zgh[g0:g1, h0:h1], M[g0:g1, h0:h1]
                if p != q:
                    W3 = M[h0:h1, g0:g1].T
                    S2[mu] += np.sum(W0 * (W1 ** 2 + W3 ** 2))
                    S1[mu] += 2.0 * np.sum(W0 * W1 * W3)
                else:
                    S2[mu] += np.sum(W0 * W1 ** 2)
                    S1[mu] += np.sum(W0 * W1 * W1.T)
    return 2.0 * np.sum(-S1 * A), 2.0 * np.sum(-2.0 * S2 * A ** 2)


def thc_1b2g_bruteforce_2g(Zgh, Xgi, Yga, U, A):
    """Ec2g = -4 sum_mu A_mu^2 sum_{aibj} U(ai) U(bj) (ab|ij) with explicit (ab|ij)."""
    # (ab|ij) = sum_{g,h} Y(g,a) Y(g,b) Z(g,h) X(h,i) X(h,j)
    abij = np.einsum("ga,gb,gh,hi,hj->abij", Yga, Yga, Zgh, Xgi, Xgi)
    e = 0.0
    for mu in range(len(A)):
        Um = U[:, :, mu]
        e += -4.0 * A[mu] ** 2 * np.einsum("ai,bj,abij->", Um, Um, abij)
    return e
