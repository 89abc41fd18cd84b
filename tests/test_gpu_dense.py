"""GPU parity of the dense FP64 building blocks (brpa_dgemm, brpa_dsyrk,
brpa_dtrsm: the library's own MFMA GEMM and triangular-solve kernels behind
the explicit-space exchange) against float64 products computed on the CPU,
on ragged shapes around the 128-wide GEMM tiles and 64-row solve blocks,
including empty inner dimensions, beta != 0 and leading dimensions larger
than the row count.  Bar: 1e-13 relative to the scale of the result (a
plain FP64 product of these sizes rounds at ~1e-15 per element)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _fmat(a, ld=None):
    """Column-major CUDA copy of the numpy matrix a (m x n), leading dimension ld."""
    m, n = a.shape
    ld = ld or m
    buf = torch.zeros((n, ld), dtype=torch.float64, device="cuda")
    buf[:, :m] = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    return buf.T[:m, :]


def _host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 5, 3), (130, 129, 16), (257, 300, 517),
                                   (64, 200, 0), (300, 64, 1000)])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_dgemm_matches_cpu(ctx, m, n, k, ta, tb):
    rng = np.random.default_rng(m * 1000 + n + 7 * k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    alpha, beta = 0.75, -1.25
    ref = alpha * (A.T if ta else A) @ (B.T if tb else B) + beta * C0
    Cd = _fmat(C0, ld=m + 3)
    ctx.dgemm(_fmat(A, ld=A.shape[0] + 1), _fmat(B), Cd, transa=ta, transb=tb, alpha=alpha,
              beta=beta)
    got = _host(Cd)
    scale = max(1.0, np.max(np.abs(ref)))
    print("max rel err %.2e" % (np.max(np.abs(got - ref)) / scale))
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale


@pytest.mark.parametrize("m,n,k", [(2, 129, 1), (130, 257, 48), (258, 300, 517), (64, 64, 2000),
                                   (300, 17, 33)])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_dgemm_even_ld_matches_cpu(ctx, m, n, k, ta, tb):
    """Products with even leading dimensions and aligned operands, which stage
    their k panels by direct global -> LDS loads (an operand contiguous in
    rows also needs an even row count; NT with odd n takes the register-staged
    form): ragged row and column tiles (rows past m, n repeat the last row or
    row pair), k not a multiple of the 16-k stage (the last stage goes through
    registers with zeros), k = 1."""
    rng = np.random.default_rng(m * 7 + n * 13 + k + 100 * ta + 1000 * tb)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    ref = 1.5 * (A.T if ta else A) @ (B.T if tb else B) - 0.5 * C0
    Cd = _fmat(C0)
    lda = A.shape[0] + A.shape[0] % 2
    ldb = B.shape[0] + B.shape[0] % 2 + 2
    ctx.dgemm(_fmat(A, ld=lda), _fmat(B, ld=ldb), Cd, transa=ta, transb=tb, alpha=1.5, beta=-0.5)
    got = _host(Cd)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale


@pytest.mark.parametrize("n,k", [(1, 5), (100, 37), (129, 300), (384, 1000)])
@pytest.mark.parametrize("trans", [False, True])
def test_dsyrk_matches_cpu(ctx, n, k, trans):
    rng = np.random.default_rng(n + 31 * k)
    A = rng.standard_normal((k, n) if trans else (n, k))
    op = A.T if trans else A
    ref = -0.5 * op @ op.T
    Cd = _fmat(np.zeros((n, n)))
    ctx.dsyrk(_fmat(A), Cd, trans=trans, alpha=-0.5, beta=0.0, fill_lower=True)
    got = _host(Cd)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale
    assert np.array_equal(got, got.T)
    # beta = 1 on an upper triangle: the strictly lower part below the diagonal
    # tiles is not touched without fill_lower
    C1 = np.triu(rng.standard_normal((n, n)))
    C1 = C1 + np.triu(C1, 1).T
    Cd = _fmat(C1)
    ctx.dsyrk(_fmat(A), Cd, trans=trans, alpha=1.0, beta=1.0, fill_lower=True)
    ref1 = op @ op.T + C1
    assert np.max(np.abs(_host(Cd) - ref1)) <= 1e-13 * np.max(np.abs(ref1))


@pytest.mark.parametrize("n,m", [(1, 3), (63, 10), (64, 64), (65, 200), (200, 7), (333, 129)])
@pytest.mark.parametrize("side_right", [False, True])
def test_dtrsm_matches_cpu(ctx, n, m, side_right):
    """B <- U^{-T} B or B U^{-1} with U the upper Cholesky factor of an SPD
    matrix whose eigenvalues span six orders of magnitude."""
    rng = np.random.default_rng(n * 17 + m)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = (Q * np.geomspace(1.0, 1e-6, n)) @ Q.T + 1e-8 * np.eye(n)
    U = np.linalg.cholesky(S).T
    B = rng.standard_normal((m, n) if side_right else (n, m))
    ref = B @ np.linalg.inv(U) if side_right else np.linalg.solve(U.T, B)
    # the solver reads only the upper triangle of U: fill the rest with noise
    Un = U + np.tril(rng.standard_normal((n, n)), -1)
    Bd = _fmat(B, ld=B.shape[0] + 2)
    ctx.dtrsm(_fmat(Un), Bd, side_right=side_right)
    got = _host(Bd)
    # backward check, independent of the conditioning of U
    resid = (got @ U - B) if side_right else (U.T @ got - B)
    print("residual %.2e  forward %.2e" % (np.max(np.abs(resid)),
                                          np.max(np.abs(got - ref)) / np.max(np.abs(ref))))
    assert np.max(np.abs(resid)) <= 1e-12 * np.max(np.abs(B))
    assert np.max(np.abs(got - ref)) <= 1e-6 * np.max(np.abs(ref))


def test_dense_rejects_bad_shapes(ctx):
    import brpa
    A = _fmat(np.ones((4, 3)))
    with pytest.raises(ValueError):
        ctx.dgemm(A, A, _fmat(np.zeros((4, 4))))
    h = ctx._h
    with pytest.raises(brpa.BrpaError):
        ctx._check(ctx._lib.brpa_dgemm(h, 0, 0, 4, 4, 3, 1.0, None, 4, None, 3, 0.0, None, 4))


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 129, 200, 333, 640, 1029])
def test_dtrtri_matches_cpu(ctx, n):
    """Upper triangular inverse (brpa_dtrtri) of the Cholesky factor of an SPD
    matrix, with noise below the diagonal that must not be read; the result has
    exact zeros below the diagonal.  Bar: |W U - 1| <= 1e-12 (W U - 1 depends on
    the conditioning only mildly for these factors) and 1e-9 relative against
    numpy's inverse."""
    rng = np.random.default_rng(n + 5)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = (Q * np.geomspace(1.0, 1e-3, n)) @ Q.T + 1e-6 * np.eye(n)
    U = np.linalg.cholesky(S).T
    Un = U + np.tril(rng.standard_normal((n, n)), -1)
    W = _fmat(np.full((n, n), np.nan))
    ctx.dtrtri(_fmat(Un, ld=n + 3), W)
    got = _host(W)
    assert np.all(np.tril(got, -1) == 0.0)
    ref = np.linalg.inv(U)
    print("|WU - 1| %.2e  rel %.2e" % (np.max(np.abs(got @ U - np.eye(n))),
                                      np.max(np.abs(got - ref)) / np.max(np.abs(ref))))
    assert np.max(np.abs(got @ U - np.eye(n))) <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_dtrtri_batched(ctx):
    """Batched inverses (strided): every matrix its own inverse."""
    rng = np.random.default_rng(3)
    nmat, n = 5, 300
    Us = [np.triu(rng.standard_normal((n, n)) * 0.05) + 2 * np.eye(n) for _ in range(nmat)]
    # C-contiguous (nmat, n, n) holding U_z column-major: [z] = U_z^T row-major
    Ud = torch.from_numpy(np.ascontiguousarray(np.stack([u.T for u in Us]))).cuda()
    Wd = torch.full_like(Ud, float("nan"))
    ctx.dtrtri(Ud, Wd)
    got = Wd.cpu().numpy()
    for z in range(nmat):
        Wz = got[z].T
        assert np.max(np.abs(Wz @ Us[z] - np.eye(n))) <= 1e-13
