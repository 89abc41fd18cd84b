"""GPU parity of the packed Pi(u) storage (brpa_pack_pi / brpa_unpack_pi) and of
E_RPA from packed Pi(u) (brpa_ecrpa_packed): the copies are bit-exact, the
packed energy equals the full-matrix energy bit for bit and the oracle's to 1e-8 Ha."""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nf", [(1, 1), (7, 3), (130, 5), (600, 2)])
def test_pack_unpack_bit_exact(ctx, n, nf):
    rng = np.random.default_rng(n)
    M = rng.standard_normal((nf, n, n))
    S = M + M.transpose(0, 2, 1)
    AP = ctx.pack_pi(S)
    assert np.array_equal(AP, erpa.pack_lower(S))
    assert np.array_equal(ctx.unpack_pi(AP, n), S)
    AP2 = ctx.pack_pi(S, packed=AP.copy(), accumulate=True)
    assert np.array_equal(AP2, AP + erpa.pack_lower(S))


def test_ecrpa_packed(ctx):
    import torch
    f = synth.make_fragment(8, 72, 400, 3)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 16)
    Pis = np.stack([erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x])
    AP = erpa.pack_lower(Pis)
    e_full = ctx.ecrpa_analytic(Pis, x, w)[0]
    e_pk, eu = ctx.ecrpa_packed(AP, 400, x, w)
    assert e_pk == e_full
    assert abs(e_pk - erpa.ecrpa_analytic(Pis, x, w)) <= 1e-8
    e_dev = ctx.ecrpa_packed(torch.from_numpy(AP).cuda(), 400, x, w)[0]
    assert e_dev == e_full
