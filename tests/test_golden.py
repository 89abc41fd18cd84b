"""Known-answer cases (tests/golden/golden.json, written by
tests/golden/make_golden.py from the oracle).

CPU: the inputs regenerate bit for bit (R hash) and the oracle still gives
the stored answers.  GPU: the library reproduces the stored answers within
the path's tolerances (1e-8 Ha FP64, 1e-7 Ha split forms) without running the
oracle.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from brpa import synth

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def _frag(c):
    f = synth.make_fragment(c["nocc"], c["nvir"], c["naux"], c["seed"])
    h = hashlib.sha256(np.ascontiguousarray(f.R).tobytes()).hexdigest()
    assert h == c["R_sha256"], "synthetic generator changed"
    return f


@pytest.mark.parametrize("c", GOLDEN["erpa"], ids=[c["name"] for c in GOLDEN["erpa"]])
def test_oracle_reproduces_golden_erpa(c):
    from oracle import erpa, grid
    f = _frag(c)
    x, w, info = grid.freq_grid(f.eocc, f.evir, c["nfreqs"])
    assert np.array_equal(x, c["freqs"]) and np.array_equal(w, c["weights"])
    e, eu = erpa.ecrpa(f.R, f.gaps(), x, w, return_terms=True)
    assert abs(e - c["ecrpa"]) <= 1e-12 * abs(c["ecrpa"])
    assert np.allclose(np.array(eu, dtype=float), c["eu"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("c", GOLDEN["t2_sosex"], ids=[c["name"] for c in GOLDEN["t2_sosex"]])
def test_oracle_reproduces_golden_t2_sosex(c):
    from oracle import erpa, grid, sosex, t2
    f = _frag(c)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, c["nfreqs"])
    g = f.gaps()
    Pis = [erpa.gpiug_mo(f.R, g, xx) for xx in x]
    A, V = t2.t2_v5(f.R, g, x, w, Pis, nguess=c["nguess"], seed=c["seed"])[:2]
    assert np.allclose(A, c["A"], rtol=1e-10, atol=1e-14)
    assert abs(sosex.sosex(f.R, V, A, c["nocc"], c["nvir"]) - c["sosex"]) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["erpa"], ids=[c["name"] for c in GOLDEN["erpa"]])
def test_library_matches_golden_erpa(ctx, c):
    import brpa
    f = _frag(c)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, c["nfreqs"])
    assert np.array_equal(x, c["freqs"]) and np.array_equal(w, c["weights"])
    for prec, tol in (("fp64", 1e-8), ("split_bf16", 1e-7), ("split_i8", 1e-7)):
        e, eu = ctx.ecrpa(f.R, f.eocc, f.evir, x, w, prec)
        assert abs(e - c["ecrpa"]) <= tol, (prec, e - c["ecrpa"])


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["t2_sosex"], ids=[c["name"] for c in GOLDEN["t2_sosex"]])
def test_library_matches_golden_t2_sosex(ctx, c):
    import brpa
    f = _frag(c)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, c["nfreqs"])
    ng = c["nguess"] or 0
    A, V = ctx.t2(f.R, f.eocc, f.evir, x, w, nguess=ng, seed=c["seed"])
    assert np.allclose(A, c["A"], rtol=1e-9, atol=1e-12)
    Tm = V @ np.diag(A) @ V.T
    assert abs(np.linalg.norm(Tm) - c["T2_frobenius"]) <= 1e-9 * c["T2_frobenius"]
    assert abs(ctx.sosex(V, A, f.R, c["nocc"], c["nvir"]) - c["sosex"]) <= 1e-8
    assert abs(ctx.sosex(V, A, f.R, c["nocc"], c["nvir"], s2_operator=True)
               - c["sosex_s2"]) <= 1e-8
