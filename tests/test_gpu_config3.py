"""The 256-fragment batched workload (o ~ U{8..40}, v ~ U{70..350},
N = round(4.5 (o+v)), 16 frequencies) on the GPU, on seeded subsets of the
batch itself (brpa.synth.batch_shapes / make_batch contents), with the grids
from the library's own host code (brpa.freq_grid).

E_RPA: fragments from all three NAux buckets of the autotuner (N < 640,
640 <= N < 1280, N >= 1280), including the batch's largest NAux (1732) and its
largest o*v (13148), in the batched launches of every form: FP64 within 1e-8
Ha of the oracle, split-BF16 and split-int8 within 1e-7 Ha.

Exchange: ten fragments of the batch, five with o*v <= NG = ceil(3 N) (the
oversampled route: the whole ov space is kept) and five with o*v > NG (three
of them keep only part of the guess space, so the Gram eigensolve decides).
They are the smallest of each kind whose Gram eigenvalues all lie at least
1.7 % (0.007 in log10) away from the 1e-10 cutoff, so that rounding cannot
flip a keep/drop decision; the next smallest fragments with o*v <= NG
(fragment 113, 0.25 %) and o*v > NG (fragment 175, 0.35 %) sit so close that
their NT2 is decided by rounding, in the reference's own eigensolver as
well.  NT2 equal to the oracle's, E_x within 1e-8 Ha (FP64) and 1e-7 Ha
(split-int8 contractions and SOSEX).
"""
import numpy as np
import pytest

import brpa
from brpa import synth
from oracle import erpa, sosex, t2

pytestmark = pytest.mark.gpu

ERPA_FRAGS = [73, 1, 34, 0, 2, 3, 4, 243, 164]
EXCH_FRAGS = [73, 28, 118, 122, 219, 128, 215, 214, 135, 188]


def _batch_frag(f):
    o, v, n = synth.batch_shapes(256)[f]
    fr = synth.make_fragment(o, v, n, synth.BASE_SEED + 1 + f)
    x, w, _ = brpa.freq_grid(fr.eocc, fr.evir, synth.BATCH_NFREQS)
    return fr, x, w


def test_config3_erpa_all_buckets(ctx):
    frags, refs = [], []
    for f in ERPA_FRAGS:
        fr, x, w = _batch_frag(f)
        Pis = [erpa.gpiug_mo(fr.R, fr.gaps(), u) for u in x]
        refs.append(erpa.ecrpa_analytic(Pis, x, w))
        frags.append((fr.R, fr.eocc, fr.evir, x, w))
    nbuckets = {0 if fr[0].shape[1] < 640 else (1 if fr[0].shape[1] < 1280 else 2)
                for fr in frags}
    assert nbuckets == {0, 1, 2}
    refs = np.array(refs)
    for prec, tol in (("fp64", 1e-8), ("split_bf16", 1e-7), ("split_i8", 1e-7)):
        e = ctx.ecrpa_batch(frags, prec)
        print("%s: max |dE| %.3e over %d fragments" % (prec, np.max(np.abs(e - refs)), len(refs)))
        assert np.max(np.abs(e - refs)) <= tol


def test_config3_exchange_both_routes(ctx):
    frags, refs, nts = [], [], []
    for f in EXCH_FRAGS:
        fr, x, w = _batch_frag(f)
        Pis = [erpa.gpiug_mo(fr.R, fr.gaps(), u) for u in x]
        A, V = t2.t2_v5(fr.R, fr.gaps(), x, w, Pis, seed=synth.BASE_SEED + f)
        refs.append(sosex.sosex(fr.R, V, A, fr.nocc, fr.nvir))
        nts.append(len(A))
        frags.append((fr.R, fr.eocc, fr.evir, x, w))
    over = [fr[0].shape[0] <= int(np.ceil(3.0 * fr[0].shape[1])) for fr in frags]
    assert sum(over) == 5
    refs, nts = np.array(refs), np.array(nts)
    # fragment f of the batch draws its guess vectors with seed BASE_SEED + f;
    # exchange_batch gives fragment k of its list seed + k, so one call per fragment
    for prec, tol in (("fp64", 1e-8), ("split_i8", 1e-7)):
        ex = np.zeros(len(EXCH_FRAGS))
        nt = np.zeros(len(EXCH_FRAGS), dtype=np.int64)
        for k, f in enumerate(EXCH_FRAGS):
            e1, n1 = ctx.exchange_batch([frags[k]], seed=synth.BASE_SEED + f, precision=prec)
            ex[k], nt[k] = e1[0], n1[0]
        print("%s: NT2 %s oracle %s; |dE_x| %s" % (prec, nt, nts, np.abs(ex - refs)))
        if prec == "fp64":
            assert np.array_equal(nt, nts)
        assert np.max(np.abs(ex - refs)) <= tol
