"""The identities behind the dropped-subspace basis of the explicit-space exchange
(beyond-rpa_amd/csrc/xspace.cpp), checked on small dense matrices in numpy
against the reference's selection rule (rpa_CC_Doubles.f90:595-641: keep the
Gram eigenvectors whose eigenvalues exceed the cutoff, T = X U s^(-1/2),
T2 = T f(T^T I T) T^T with f(a) = a / (2 (1 + a))).

- oversampled (G = X X^T): with Zd the eigenvectors of G at or below the cutoff
  and P = 1 - Zd Zd^T, T2 = (1 - (1 + P I P)^(-1)) / 2;
- otherwise (G = X^T X): Cholesky-QR of the padded matrix [X P; Zd^T] gives a
  basis whose top K rows T' satisfy T2 = (T' T'^T - W W^T) / 2, W = T' U^(-1),
  U^T U = 1 + T'^T I T';
- the count certificate: exactly md eigenvalues lie at or below the cutoff when
  the compression P G P + Zd Zd^T - cutoff has a Cholesky factor and the md Ritz
  values of Zd lie at or below it.
"""
import numpy as np

CUT = 1e-10


def _reference_t2(X, I, over):
    """T2 by the reference's rule: the Gram eigenvectors of X^T X with eigenvalues
    above the cutoff span the kept subspace, here taken as the left singular
    vectors of X with s^2 > cutoff (the same subspace, exactly orthonormal)."""
    Ul, sv, _ = np.linalg.svd(X, full_matrices=False)
    keep = sv ** 2 > CUT
    T = Ul[:, keep]
    a, E = np.linalg.eigh(T.T @ I @ T)
    V = T @ E
    return (V * (0.5 * a / (1.0 + a))) @ V.T, int(keep.sum())


def _problem(K, NG, ndrop, seed):
    """I negative semidefinite (K x K) and X (K x NG) whose Gram matrix has ndrop
    eigenvalues well below the cutoff, the rest well above it."""
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((3 * K, K)) / np.sqrt(3 * K)
    I = -0.25 * (Y.T @ Y)                                     # 1 + I positive definite
    r = min(K, NG)
    Ul, _ = np.linalg.qr(rng.standard_normal((K, r)))
    Vr, _ = np.linalg.qr(rng.standard_normal((NG, r)))
    sv = np.exp(rng.uniform(np.log(1e-3), np.log(3.0), r))
    sv[:ndrop] = np.sqrt(CUT) * rng.uniform(1e-3, 0.3, ndrop)   # Gram eigenvalues <= 0.09 cutoff
    X = (Ul * sv) @ Vr.T
    return X, I


def _dropped(G, md):
    w, Z = np.linalg.eigh(G)
    assert np.sum(w <= CUT) == md
    return Z[:, :md], w


def test_projected_form_oversampled():
    K, NG, md = 40, 70, 5
    X, I = _problem(K, NG, md, 1)
    ref, nt = _reference_t2(X, I, True)
    assert nt == K - md
    Zd, _ = _dropped(X @ X.T, md)
    P = np.eye(K) - Zd @ Zd.T
    # P I P as the library forms it: I - Zd E^T - E Zd^T, E = I Zd - Zd (Zd^T I Zd) / 2
    E = I @ Zd - 0.5 * Zd @ (Zd.T @ I @ Zd)
    PIP = I - Zd @ E.T - E @ Zd.T
    assert np.allclose(PIP, P @ I @ P, atol=1e-13)
    t2 = 0.5 * (np.eye(K) - np.linalg.inv(np.eye(K) + PIP))
    # the identity is exact; 1e-9 covers the angle of the computed Zd (eigh's
    # absolute error ~ eps |G| over the spectral gap ~1e-6); a wrong form is off by O(1)
    assert np.max(np.abs(t2 - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_padded_cholesky_qr_form():
    K, NG, md = 60, 35, 4
    X, I = _problem(K, NG, md, 2)
    ref, nt = _reference_t2(X, I, False)
    assert nt == NG - md
    Zd, _ = _dropped(X.T @ X, md)
    Xp = X - (X @ Zd) @ Zd.T
    B = Zd.T.copy()
    # pass 1 with the shifted factor, pass 2 with the Gram matrix of both blocks
    G1 = Xp.T @ Xp + B.T @ B - CUT * np.eye(NG)
    U1 = np.linalg.cholesky(G1).T
    Xp = np.linalg.solve(U1.T, Xp.T).T
    B = np.linalg.solve(U1.T, B.T).T
    U2 = np.linalg.cholesky(Xp.T @ Xp + B.T @ B).T
    Tt = np.linalg.solve(U2.T, Xp.T).T                    # top K rows of the padded basis
    U = np.linalg.cholesky(np.eye(NG) + Tt.T @ I @ Tt).T
    W = np.linalg.solve(U.T, Tt.T).T
    t2 = 0.5 * (Tt @ Tt.T - W @ W.T)
    assert np.max(np.abs(t2 - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_count_certificate():
    K, NG, md = 30, 50, 3
    X, _ = _problem(K, NG, md, 3)
    G = X @ X.T
    Zd, w = _dropped(G, md)
    P = np.eye(K) - Zd @ Zd.T
    D = G @ Zd - 0.5 * Zd @ (Zd.T @ G @ Zd + np.eye(md))
    C = G - Zd @ D.T - D @ Zd.T - CUT * np.eye(K)
    assert np.allclose(C, P @ G @ P + Zd @ Zd.T - CUT * np.eye(K), atol=1e-13)
    np.linalg.cholesky(C)                                 # the complement clears the cutoff
    # a subspace that misses one dropped direction leaves the compression indefinite
    Zm = Zd[:, :md - 1]
    Pm = np.eye(K) - Zm @ Zm.T
    Cm = Pm @ G @ Pm + Zm @ Zm.T - CUT * np.eye(K)
    assert np.min(np.linalg.eigvalsh(Cm)) < 0.0
