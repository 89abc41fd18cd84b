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


def _h(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()


def _thc_frag(c):
    f = synth.make_thc_fragment(c["nocc"], c["nvir"], c["naux"], c["ngrid"], c["seed"])
    assert _h(f.Zgk) == c["Z_sha256"], "synthetic generator changed"
    return f


def _aomo_case(c):
    from oracle import aomo
    sizes = c["shell_sizes"]
    loc = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = aomo.pair_table([(A, B) for A in range(len(sizes)) for B in range(A + 1)], loc, sizes)
    rng = np.random.default_rng(c["seed"])
    nao = int(sum(sizes))
    RG = rng.standard_normal((len(pairs), c["nvecs"]))
    assert _h(RG) == c["RG_sha256"]
    Co = rng.standard_normal((c["nocc"], nao))
    Cv = rng.standard_normal((c["nvir"], nao))
    return pairs, RG, Co, Cv


def _cut_amplitudes(c, library=None):
    """(f, x, w, A, V) for a t2_cutoff_mbpt3 case, from the oracle or the library."""
    from oracle import erpa, grid, t2
    f = _frag(c)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, c["nfreqs"])
    if library is not None:
        A, V = library.t2(f.R, f.eocc, f.evir, x, w, seed=c["seed"])
        return f, x, w, A, V
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A, V = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=c["seed"])
    return f, x, w, A, V


@pytest.mark.parametrize("c", GOLDEN["thc"], ids=[c["name"] for c in GOLDEN["thc"]])
def test_oracle_reproduces_golden_thc(c):
    from oracle import erpa, grid, thc
    f = _thc_frag(c)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, c["nfreqs"])
    Pis, R = thc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, x)
    assert abs(np.linalg.norm(R) - c["R_frobenius"]) <= 1e-13 * c["R_frobenius"]
    assert np.allclose([np.trace(P) for P in Pis], c["pi_trace"], rtol=1e-13, atol=0)
    assert abs(erpa.ecrpa_analytic(Pis, x, w) - c["ecrpa"]) <= 1e-12 * abs(c["ecrpa"])


@pytest.mark.parametrize("c", GOLDEN["t2_cutoff_mbpt3"],
                         ids=[c["name"] for c in GOLDEN["t2_cutoff_mbpt3"]])
def test_oracle_reproduces_golden_cutoff_mbpt3(c):
    from oracle import mbpt3, t2
    f, x, w, A, V = _cut_amplitudes(c)
    Ac, Vc, K, _, nsc = t2.t2_cutoff(A, V, c["cutoff_type"], c["thresh"], c["nocc"],
                                     smooth_step_on=c["smooth"])
    assert len(Ac) == c["nvecs_kept"] and nsc == c["n_scaled"]
    assert np.allclose(Ac, c["A_kept"], rtol=1e-10, atol=1e-14)
    o, v = c["nocc"], c["nvir"]
    assert abs(mbpt3.mbpt3_1(f.R, V, A, o, v, f.eocc, f.evir) - c["mbpt3_1"]) <= 1e-10 * abs(c["mbpt3_1"])
    assert abs(mbpt3.mbpt3_2(f.R, V, A, o, v) - c["mbpt3_2"]) <= 1e-10 * abs(c["mbpt3_2"])


@pytest.mark.parametrize("c", GOLDEN["aomo"], ids=[c["name"] for c in GOLDEN["aomo"]])
def test_oracle_reproduces_golden_aomo(c):
    from oracle import aomo
    pairs, RG, Co, Cv = _aomo_case(c)
    R = aomo.crg(RG, pairs, Co.T, Cv.T)
    assert abs(np.linalg.norm(R) - c["CRG_frobenius"]) <= 1e-13 * c["CRG_frobenius"]


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["thc"], ids=[c["name"] for c in GOLDEN["thc"]])
def test_library_matches_golden_thc(ctx, c):
    f = _thc_frag(c)
    import brpa
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, c["nfreqs"])
    Pi, R = ctx.thc_piu(f.Xga, f.Xgi, f.Zgk, f.eocc, f.evir, x, want_R=True)
    assert abs(np.linalg.norm(R) - c["R_frobenius"]) <= 1e-12 * c["R_frobenius"]
    assert abs(ctx.ecrpa_analytic(Pi, x, w)[0] - c["ecrpa"]) <= 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["t2_cutoff_mbpt3"],
                         ids=[c["name"] for c in GOLDEN["t2_cutoff_mbpt3"]])
def test_library_matches_golden_cutoff_mbpt3(ctx, c):
    f, x, w, A, V = _cut_amplitudes(c, library=ctx)
    Ac, Vt, eig, common, nsc = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, seed=c["seed"],
                                          t2_cutoff_thresh=c["thresh"],
                                          t2_cutoff_type=c["cutoff_type"],
                                          smooth_step=c["smooth"])
    assert len(Ac) == c["nvecs_kept"] and nsc == c["n_scaled"]
    assert np.allclose(Ac, c["A_kept"], rtol=1e-9, atol=1e-12)
    assert abs(eig.sum() - c["eig_rpa_sum"]) <= 1e-9 * abs(c["eig_rpa_sum"])
    o, v = c["nocc"], c["nvir"]
    assert abs(ctx.mbpt3_1(V, A, f.R, f.eocc, f.evir) - c["mbpt3_1"]) <= 1e-8 * abs(c["mbpt3_1"])
    assert abs(ctx.mbpt3_2(f.R, V, A, o, v) - c["mbpt3_2"]) <= 1e-8 * abs(c["mbpt3_2"])


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["thc_1b2g"], ids=[c["name"] for c in GOLDEN["thc_1b2g"]])
def test_library_matches_golden_thc_1b2g(ctx, c):
    import brpa
    f = _thc_frag(c)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, c["nfreqs"])
    _, R = ctx.thc_piu(f.Xga, f.Xgi, f.Zgk, f.eocc, f.evir, x, want_R=True)
    A, V = ctx.t2(R, f.eocc, f.evir, x, w, seed=c["seed"])
    assert len(A) == c["nvecs_t2"]
    Uaim = np.ascontiguousarray(V.T)
    Zgh = np.ascontiguousarray(f.Zgk.T @ f.Zgk)
    e1, e2, _, _, _ = ctx.thc_1b2g(Zgh, f.Xgi, f.Xga, Uaim, A)
    assert abs(e1 - c["ec1b"]) <= 1e-8 * abs(c["ec1b"])
    assert abs(e2 - c["ec2g"]) <= 1e-8 * abs(c["ec2g"])


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["aomo"], ids=[c["name"] for c in GOLDEN["aomo"]])
def test_library_matches_golden_aomo(ctx, c):
    pairs, RG, Co, Cv = _aomo_case(c)
    R = ctx.ao_mo(RG, pairs, Co, Cv)
    wts = np.cos(np.arange(R.size)).reshape(R.shape)
    assert abs(np.linalg.norm(R) - c["CRG_frobenius"]) <= 1e-12 * c["CRG_frobenius"]
    assert abs(np.sum(R * wts) - c["CRG_weighted_sum"]) <= 1e-10 * c["CRG_frobenius"]


@pytest.mark.gpu
@pytest.mark.parametrize("c", GOLDEN["rpa_g"], ids=[c["name"] for c in GOLDEN["rpa_g"]])
def test_library_matches_golden_rpa_g(ctx, c):
    import brpa
    f = _frag(c)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, c["nfreqs"])
    RWR = np.zeros((c["naux"], c["naux"]))
    ctx.gpiug_mo(RWR, f.R, f.eocc, f.evir, x[0])
    G, R2, nv, err = ctx.rpa_g(RWR, "eigen", c["target"])
    assert nv == c["nvecs"]
    assert np.allclose(np.diag(R2)[:8], c["eigenvalues_top"], rtol=1e-11, atol=0)
