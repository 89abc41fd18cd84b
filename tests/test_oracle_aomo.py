"""Pins of the AO -> MO transform oracle (rpa_CRG / rpa_C), CPU only."""
import numpy as np

from oracle import aomo

SIZES = [1, 3, 5, 1, 3, 7]       # s, p, d, s, p, f-like shells


def _shells():
    loc = np.concatenate([[0], np.cumsum(SIZES)[:-1]])
    pairs_sh = [(A, B) for A in range(len(SIZES)) for B in range(A + 1)]
    return loc, pairs_sh


def test_pair_table_covers_every_unordered_pair_once():
    loc, sp = _shells()
    nao = sum(SIZES)
    P = aomo.pair_table(sp, loc, SIZES)
    assert len(P) == nao * (nao + 1) // 2
    seen = {(max(p, q), min(p, q)) for p, q in P}
    assert len(seen) == len(P)


def test_packed_equals_unpacked_contraction():
    loc, sp = _shells()
    nao = sum(SIZES)
    pairs = aomo.pair_table(sp, loc, SIZES)
    rng = np.random.default_rng(1)
    RG = rng.standard_normal((len(pairs), 9))
    Co = rng.standard_normal((nao, 3))
    Cv = rng.standard_normal((nao, 4))
    R = aomo.crg(RG, pairs, Co, Cv)
    assert np.max(np.abs(R - aomo.crg_unpacked(RG, pairs, Co, Cv))) < 1e-12


def test_planted_missing_exchange_term_fails():
    """Dropping the Co(q, i) Cv(p, a) term (a plausible slip) breaks the identity."""
    loc, sp = _shells()
    nao = sum(SIZES)
    pairs = aomo.pair_table(sp, loc, SIZES)
    rng = np.random.default_rng(2)
    RG = rng.standard_normal((len(pairs), 5))
    Co = rng.standard_normal((nao, 2))
    Cv = rng.standard_normal((nao, 3))
    p, q = pairs[:, 0], pairs[:, 1]
    wrong = (Co[p, :].T[:, None, :] * Cv[q, :].T[None, :, :]).reshape(6, -1) @ RG
    assert np.max(np.abs(wrong - aomo.crg_unpacked(RG, pairs, Co, Cv))) > 1e-3
