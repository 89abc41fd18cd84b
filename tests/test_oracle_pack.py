"""Pins of the packed Pi(u) storage oracle against LAPACK's dtrttp / dtpttr,
the routines the reference calls (real_linalg.f90:749, 772)."""
import numpy as np
import scipy.linalg.lapack as lapack

from oracle import erpa


def test_pack_matches_dtrttp_and_round_trips():
    rng = np.random.default_rng(5)
    n = 9
    M = rng.standard_normal((3, n, n))
    S = M + M.transpose(0, 2, 1)
    AP = erpa.pack_lower(S)
    APm = erpa.pack_lower(M)      # non-symmetric: which triangle is read matters
    for w in range(3):
        # our [w][j][i] is Fortran column-major A(i, j): pass the Fortran array A = M[w].T
        refm, info = lapack.dtrttp(np.asfortranarray(M[w].T), uplo="L")
        assert info == 0 and np.array_equal(APm[w], refm)
        ref, info = lapack.dtrttp(np.asfortranarray(S[w].T), uplo="L")
        assert info == 0 and np.array_equal(AP[w], ref)
        full, info = lapack.dtpttr(n, ref, uplo="L")
        assert info == 0
        assert np.array_equal(np.tril(full), np.tril(S[w].T))
    assert np.array_equal(erpa.unpack_lower(AP, n), S)
