"""GPU parity of the allocation-free entry point brpa_ecrpa_analytic_noalloc
against brpa_ecrpa_analytic, through the C ABI: two calls in a row with the
second one smaller (the scratch array keeps the first call's contents), the
frequency-chunked path (buffers larger than the scratch array), device inputs,
and the size limit.  The bar is FP64 parity: |dE| <= 1e-8 Ha against the
allocating entry point and against the oracle."""
import numpy as np
import pytest

import brpa
from brpa import synth
from oracle import erpa, grid

pytestmark = pytest.mark.gpu

SCRATCH = 256 << 20   # BRPA_NOALLOC_SCRATCH_MB


def _host_case(o, v, n, nf, seed):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis = np.stack([erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x])
    return Pis, x, w


def test_noalloc_two_calls_second_smaller():
    with brpa.Context(0) as ctx:
        for (o, v, n, nf, seed) in ((8, 72, 400, 16, 11), (4, 30, 130, 5, 12)):
            print("case o=%d v=%d N=%d F=%d" % (o, v, n, nf), flush=True)
            Pis, x, w = _host_case(o, v, n, nf, seed)
            e_ref, eu_ref = ctx.ecrpa_analytic(Pis, x, w)
            e, eu = ctx.ecrpa_analytic_noalloc(Pis, x, w)
            e_or = erpa.ecrpa_analytic(list(Pis), x, w)
            print("  E %.12f  dE vs allocating %.2e  vs oracle %.2e  bit-equal %s" %
                  (e, e - e_ref, e - e_or, e == e_ref), flush=True)
            assert abs(e - e_ref) <= 1e-8
            assert abs(e - e_or) <= 1e-8
            assert np.max(np.abs(eu - eu_ref)) <= 1e-8


def test_noalloc_chunked_device_inputs():
    import torch
    dev = torch.device("cuda:0")
    with brpa.Context(0) as ctx:
        # N = 2400: 46 MB per frequency, so the 16 frequencies run in chunks;
        # then a smaller call that fits in one chunk
        for (n, nf, m, seed) in ((2400, 16, 96, 21), (1200, 3, 64, 22)):
            print("case N=%d F=%d (%d chunk bytes per frequency)" % (n, nf, 8 * n * n),
                  flush=True)
            g = torch.Generator(device="cpu").manual_seed(seed)
            B = (0.05 * torch.randn(nf, n, m, generator=g, dtype=torch.float64)).to(dev)
            Pi = torch.bmm(B, B.transpose(1, 2).contiguous()).contiguous()
            x = np.linspace(0.1, 3.0, nf)
            w = np.linspace(0.5, 0.1, nf)
            torch.cuda.synchronize()
            e_ref, eu_ref = ctx.ecrpa_analytic(Pi, x, w)
            e, eu = ctx.ecrpa_analytic_noalloc(Pi, x, w)
            # the same energy in float64 on the CPU: sum_u w/(2 pi) (logdet(1+Pi) - tr Pi)
            P = Pi.cpu().numpy()
            e_cpu = sum(w[k] / (2 * np.pi) * (np.linalg.slogdet(np.eye(n) + P[k])[1] -
                                              np.trace(P[k])) for k in range(nf))
            print("  E %.12f  dE vs allocating %.2e  vs numpy %.2e" % (e, e - e_ref, e - e_cpu),
                  flush=True)
            assert abs(e - e_ref) <= 1e-8
            assert abs(e - e_cpu) <= 1e-8
            assert np.max(np.abs(eu - eu_ref)) <= 1e-8


def test_noalloc_rejects_beyond_scratch():
    n = 6000   # 288 MB for one frequency > the scratch array
    assert 8 * n * n > SCRATCH
    Pi = np.zeros((1, n, n))
    with brpa.Context(0) as ctx:
        with pytest.raises(brpa.BrpaError, match="scratch"):
            ctx.ecrpa_analytic_noalloc(Pi, np.array([1.0]), np.array([1.0]))
