"""GPU parity of the E_RPA hot path against the CPU oracle (through the C ABI).

Tolerances: 1e-8 Ha for the FP64 form, 1e-7 Ha for the split forms (bf16 and
int8 slices; absolute, per fragment), relative 1e-12 for FP64 Pi(u) elements.
"""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-8
TOL_SPLIT = 1e-7
SPLITS = ("split_bf16", "split_i8")


def _frag_and_grid(o, v, n, nf, seed):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    return f, x, w


@pytest.mark.parametrize("o,v,n", [(4, 20, 100), (3, 7, 5), (2, 50, 130), (1, 17, 64)])
def test_gpiug_mo_matches_oracle(ctx, o, v, n):
    f = synth.make_fragment(o, v, n, 11)
    g = f.gaps()
    P0 = np.random.default_rng(1).standard_normal((n, n))
    P0 = P0 + P0.T
    ref = P0 + erpa.gpiug_mo(f.R, g, 0.37, 4.0)
    out = ctx.gpiug_mo(P0.copy(), f.R, f.eocc, f.evir, 0.37, 4.0)
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
    # open-shell prefactor 2 (rpa_core_MO.f90:379-383)
    out2 = ctx.gpiug_mo(np.zeros((n, n)), f.R, f.eocc, f.evir, 1.3, 2.0)
    assert np.allclose(out2, erpa.gpiug_mo(f.R, g, 1.3, 2.0), rtol=1e-12, atol=1e-14)


def test_ecrpa_analytic_matches_oracle(ctx):
    f, x, w = _frag_and_grid(5, 40, 150, 12, 3)
    Pis = np.stack([erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x])
    e_ref, eu_ref = erpa.ecrpa_analytic(list(Pis), x, w, return_terms=True)
    e, eu = ctx.ecrpa_analytic(np.ascontiguousarray(Pis), x, w)
    assert abs(e - e_ref) <= TOL_FP64
    assert np.max(np.abs(eu - np.array(eu_ref))) <= 1e-8


@pytest.mark.parametrize("o,v,n,nf", [(8, 72, 400, 16), (3, 11, 70, 8), (6, 45, 257, 10),
                                      (2, 9, 3, 8)])
def test_ecrpa_fp64_matches_oracle(ctx, o, v, n, nf):
    f, x, w = _frag_and_grid(o, v, n, nf, 20261202)
    e_ref, eu_ref = erpa.ecrpa(f.R, f.gaps(), x, w, return_terms=True)
    e, eu = ctx.ecrpa(f.R, f.eocc, f.evir, x, w, "fp64")
    print("fp64 o=%d v=%d n=%d: E=%.12f ref=%.12f dE=%.2e" % (o, v, n, e, e_ref, e - e_ref))
    assert abs(e - e_ref) <= TOL_FP64
    assert np.max(np.abs(eu - np.array(eu_ref))) <= 1e-8


@pytest.mark.parametrize("prec", SPLITS)
@pytest.mark.parametrize("o,v,n,nf", [(8, 72, 400, 16), (3, 11, 70, 8), (6, 45, 257, 10),
                                      (12, 100, 500, 16), (1, 3, 200, 8), (2, 9, 3, 8)])
def test_ecrpa_split_matches_oracle(ctx, prec, o, v, n, nf):
    f, x, w = _frag_and_grid(o, v, n, nf, 20261202)
    e_ref = erpa.ecrpa(f.R, f.gaps(), x, w)
    e, _ = ctx.ecrpa(f.R, f.eocc, f.evir, x, w, prec)
    print("%s o=%d v=%d n=%d: E=%.12f ref=%.12f dE=%.2e" % (prec, o, v, n, e, e_ref, e - e_ref))
    assert abs(e - e_ref) <= TOL_SPLIT


def test_ecrpa_single_large_fragment_both_splits(ctx):
    """o=36, v=300, N=1800, 24 frequencies: both split forms within 1e-7 Ha
    of the oracle on the single large fragment."""
    f, x, w = _frag_and_grid(36, 300, 1800, 24, 20261202)
    e_ref = erpa.ecrpa(f.R, f.gaps(), x, w)
    for prec in SPLITS:
        e, _ = ctx.ecrpa(f.R, f.eocc, f.evir, x, w, prec)
        print("%s large: E=%.10f dE=%.2e" % (prec, e, e - e_ref))
        assert abs(e - e_ref) <= TOL_SPLIT


def test_split_i8_rejects_k_beyond_exact_range(ctx):
    import brpa
    o, v, n = 2, 40000, 4
    f = synth.make_fragment(o, v, n, 3)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 8)
    with pytest.raises(brpa.BrpaError) as ei:
        ctx.ecrpa(f.R, f.eocc, f.evir, x, w, "split_i8")
    assert ei.value.code == brpa._lib.ERR_ARG


def test_batch_uneven_fragments(ctx):
    shapes = [(3, 20, 60), (8, 72, 400), (5, 33, 129), (2, 5, 7), (7, 61, 300)]
    frags, refs = [], []
    for i, (o, v, n) in enumerate(shapes):
        f, x, w = _frag_and_grid(o, v, n, 8 + 2 * i, 100 + i)
        frags.append((f.R, f.eocc, f.evir, x, w))
        refs.append(erpa.ecrpa(f.R, f.gaps(), x, w))
    e64 = ctx.ecrpa_batch(frags, "fp64")
    print("batch dE fp64", e64 - refs)
    assert np.max(np.abs(e64 - refs)) <= TOL_FP64
    for prec in SPLITS:
        es = ctx.ecrpa_batch(frags, prec)
        print("batch dE", prec, es - refs)
        assert np.max(np.abs(es - refs)) <= TOL_SPLIT
    # the batched result equals the single-fragment call bit for bit
    e_single = [ctx.ecrpa(*fr, precision="fp64")[0] for fr in frags]
    assert np.array_equal(e64, np.array(e_single))


def test_device_resident_inputs(ctx):
    torch = pytest.importorskip("torch")
    f, x, w = _frag_and_grid(6, 40, 200, 12, 5)
    Rd = torch.from_numpy(f.R).cuda()
    e_dev, _ = ctx.ecrpa(Rd, f.eocc, f.evir, x, w, "fp64")
    e_host, _ = ctx.ecrpa(f.R, f.eocc, f.evir, x, w, "fp64")
    assert e_dev == e_host


def test_deterministic(ctx):
    f, x, w = _frag_and_grid(5, 50, 300, 12, 9)
    a = [ctx.ecrpa(f.R, f.eocc, f.evir, x, w, p)[0]
         for p in ("fp64", "fp64", "split_bf16", "split_bf16", "split_i8", "split_i8")]
    assert a[0] == a[1] and a[2] == a[3] and a[4] == a[5]


def test_error_codes(ctx):
    import brpa
    f, x, w = _frag_and_grid(3, 10, 20, 8, 1)
    bad = f.eocc.copy()
    bad[-1] = f.evir[0] + 0.1   # non-positive gap
    with pytest.raises(brpa.BrpaError) as ei:
        ctx.ecrpa(f.R, bad, f.evir, x, w)
    assert ei.value.code == brpa._lib.ERR_GAP
    nan = f.evir.copy()
    nan[3] = np.nan
    with pytest.raises(brpa.BrpaError) as ei:
        ctx.ecrpa(f.R, f.eocc, nan, x, w)
    assert ei.value.code == brpa._lib.ERR_NONFINITE
    # 1 + Pi not positive definite: a negative "Pi" handed to the analytic form
    P = -3.0 * np.eye(20)[None].repeat(2, axis=0)
    with pytest.raises(brpa.BrpaError) as ei:
        ctx.ecrpa_analytic(np.ascontiguousarray(P), x[:2], w[:2])
    assert ei.value.code == brpa._lib.ERR_NOT_SPD


def test_batch_many_fragments_cholesky_steps(ctx):
    """48 fragments with 5-10 Cholesky steps each in one batch: every
    work-group of a step must see the unfactored diagonal block (regression
    test for a read-after-write race between work-groups of one launch)."""
    rng = np.random.default_rng(7)
    frags, refs = [], []
    for i in range(48):
        o, v = int(rng.integers(3, 9)), int(rng.integers(30, 80))
        n = int(rng.integers(300, 640))
        f, x, w = _frag_and_grid(o, v, n, 8, 500 + i)
        frags.append((f.R, f.eocc, f.evir, x, w))
        refs.append(erpa.ecrpa(f.R, f.gaps(), x, w))
    for prec, tol in (("fp64", TOL_FP64), ("split_bf16", TOL_SPLIT), ("split_i8", TOL_SPLIT)):
        e = ctx.ecrpa_batch(frags, prec)
        print(prec, "max |dE|", np.max(np.abs(e - refs)))
        assert np.max(np.abs(e - refs)) <= tol
