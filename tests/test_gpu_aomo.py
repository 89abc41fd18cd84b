"""GPU parity of the AO -> MO transform (brpa_ao_mo) against the CPU oracle
(oracle/aomo.py).  FP64 on both sides: within 1e-13 of the largest element."""
import numpy as np
import pytest

from oracle import aomo

pytestmark = pytest.mark.gpu


def _case(sizes, nocc, nvir, nvecs, seed):
    loc = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
    sp = [(A, B) for A in range(len(sizes)) for B in range(A + 1)]
    pairs = aomo.pair_table(sp, loc, sizes)
    nao = int(sum(sizes))
    rng = np.random.default_rng(seed)
    RG = rng.standard_normal((len(pairs), nvecs))
    Co = rng.standard_normal((nocc, nao))
    Cv = rng.standard_normal((nvir, nao))
    return pairs, RG, Co, Cv


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("sizes,nocc,nvir,nvecs", [([1], 1, 1, 1), ([1, 3, 5], 2, 3, 7),
                                                   ([1, 3, 5, 1, 3, 7] * 3, 5, 40, 150),
                                                   ([1, 1, 3, 3, 5, 7] * 12, 12, 130, 700)])
def test_ao_mo_matches_oracle(ctx, sizes, nocc, nvir, nvecs):
    pairs, RG, Co, Cv = _case(sizes, nocc, nvir, nvecs, len(sizes))
    ref = aomo.crg(RG, pairs, Co.T, Cv.T)
    R = ctx.ao_mo(RG, pairs, Co, Cv)
    print("npq", len(pairs), "rel", _rel(R, ref))
    assert _rel(R, ref) <= 1e-13


def test_ao_mo_subsets_accumulate(ctx):
    """Two subsets of AO pairs accumulated into one CRG (rpa_CRG's subset loop)."""
    pairs, RG, Co, Cv = _case([1, 3, 5, 1, 3, 7] * 4, 6, 50, 200, 3)
    h = len(pairs) // 3
    R = ctx.ao_mo(RG[:h], pairs[:h], Co, Cv)
    R = ctx.ao_mo(np.ascontiguousarray(RG[h:]), pairs[h:], Co, Cv, out=R, accumulate=True)
    assert _rel(R, aomo.crg(RG, pairs, Co.T, Cv.T)) <= 1e-13


def test_ao_mo_device_tensors_and_bad_pairs(ctx):
    import torch

    import brpa
    pairs, RG, Co, Cv = _case([1, 3, 5, 7] * 5, 4, 30, 128, 4)
    d = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    R = ctx.ao_mo(d(RG), pairs, d(Co), d(Cv))
    assert R.is_cuda and _rel(R.cpu().numpy(), aomo.crg(RG, pairs, Co.T, Cv.T)) <= 1e-13
    bad = pairs.copy()
    bad[3, 1] = Co.shape[1]
    with pytest.raises(brpa.BrpaError):
        ctx.ao_mo(RG, bad, Co, Cv)
