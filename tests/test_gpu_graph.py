"""hipGraph replay of the batched E_RPA launch sequence (brpa_ctx_set_graphs).

Every replay is compared bit for bit with the eager (non-graph) result of the
same inputs, with eager library work, torch work and synchronisations between
the replays, and with new input values of the same shapes (the replay must
read the refreshed staging, not the captured values).
"""
import numpy as np
import pytest

from brpa import synth
from oracle import erpa, grid

pytestmark = pytest.mark.gpu

SHAPES = [(3, 20, 60), (8, 72, 400), (5, 33, 129), (2, 5, 7), (7, 61, 300)]


def _batch(torch, scale=1.0, seed0=100):
    out, refs = [], []
    for i, (o, v, n) in enumerate(SHAPES):
        f = synth.make_fragment(o, v, n, seed0 + i)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, 8 + 2 * i)
        R = f.R * scale
        out.append((torch.from_numpy(R).cuda(), f.eocc, f.evir, x, w))
        refs.append((R, f.gaps(), x, w))
    return out, refs


@pytest.mark.parametrize("prec", ["fp64", "split_bf16", "split_i8"])
def test_graph_replays_match_eager(ctx, prec):
    torch = pytest.importorskip("torch")
    frags, refs = _batch(torch)
    frags2, refs2 = _batch(torch, scale=1.05, seed0=300)
    ctx.set_graphs(False)
    eager = ctx.ecrpa_batch(frags, prec)
    eager2 = ctx.ecrpa_batch(frags2, prec)
    ctx.set_graphs(True)
    try:
        for rep in range(4):
            e = ctx.ecrpa_batch(frags, prec)
            cap, rpl = ctx.graph_stats()
            assert (cap, rpl) == ((1, 0) if rep == 0 else (0, 1))
            assert np.array_equal(e, eager), (rep, e - eager)
            # eager work between replays: a host-R call (not graphed) and torch work
            f = frags[1]
            ctx.ecrpa(f[0].cpu().numpy(), f[1], f[2], f[3], f[4], prec)
            t = torch.randn(512, 512, device="cuda", dtype=torch.float64)
            (t @ t).sum().item()
            torch.cuda.synchronize()
            # same shapes, new values: replayed, reads the new staging
            e2 = ctx.ecrpa_batch(frags2, prec)
            assert ctx.graph_stats() == (0, 1)
            assert np.array_equal(e2, eager2), (rep, e2 - eager2)
    finally:
        ctx.set_graphs(False)
    tol = 1e-8 if prec == "fp64" else 1e-7
    for (R, g, x, w), ev in zip(refs2, eager2):
        assert abs(ev - erpa.ecrpa(R, g, x, w)) <= tol


def test_graph_not_used_for_host_inputs(ctx):
    f = synth.make_fragment(4, 30, 100, 5)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 8)
    ctx.set_graphs(True)
    try:
        e = ctx.ecrpa_batch([(f.R, f.eocc, f.evir, x, w)], "fp64")
        assert ctx.graph_stats() == (0, 0)
    finally:
        ctx.set_graphs(False)
    assert abs(e[0] - erpa.ecrpa(f.R, f.gaps(), x, w)) <= 1e-8


def test_autotune_chooses_and_agrees(ctx):
    """Every split-int8 product variant runs on fragments of all three NAux
    buckets (the tuner checks that they agree bit for bit); the tuned batch
    stays within 1e-7 Ha of the oracle."""
    shapes = [(4, 60, 300), (9, 80, 700), (6, 250, 1300), (3, 40, 200)]
    frags, refs = [], []
    for i, (o, v, n) in enumerate(shapes):
        f = synth.make_fragment(o, v, n, 700 + i)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, 8)
        frags.append((f.R, f.eocc, f.evir, x, w))
        refs.append(erpa.ecrpa(f.R, f.gaps(), x, w))
    res = ctx.autotune(frags, "split_i8")
    print("autotune", res)
    assert sorted(res) == [0, 1, 2]
    for b, (v, ms) in res.items():
        assert 0 <= v < 3 and all(m > 0 for m in ms)
    e = ctx.ecrpa_batch(frags, "split_i8")
    assert np.max(np.abs(e - np.array(refs))) <= 1e-7


@pytest.mark.skip(reason="the persistent Cholesky launch did not complete on the GPU (twice), "
                         "cause unknown; the kernel was taken out of the library and the batch "
                         "runs as graph-replayed per-step launches")
@pytest.mark.parametrize("prec", ["fp64", "split_i8"])
def test_persistent_cholesky_matches_per_step(ctx, prec, monkeypatch):
    """The one-launch persistent Cholesky (dependency counters between work
    items) gives the same energies, bit for bit, as one launch per step, on
    uneven fragments with 1..12 Cholesky steps; also under graph replay."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    frags = []
    for i in range(24):
        o, v = int(rng.integers(2, 9)), int(rng.integers(10, 90))
        n = int(rng.integers(20, 760))
        f = synth.make_fragment(o, v, n, 1200 + i)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, 6 + i % 5)
        frags.append((torch.from_numpy(f.R).cuda(), f.eocc, f.evir, x, w))
    monkeypatch.setenv("BRPA_CHOL_PERSISTENT", "0")
    ref = ctx.ecrpa_batch(frags, prec)
    monkeypatch.setenv("BRPA_CHOL_PERSISTENT", "1")
    try:
        e = ctx.ecrpa_batch(frags, prec)
        assert np.array_equal(e, ref), e - ref
        ctx.set_graphs(True)
        for _ in range(3):
            e = ctx.ecrpa_batch(frags, prec)
            assert np.array_equal(e, ref), e - ref
    finally:
        ctx.set_graphs(False)
        monkeypatch.setenv("BRPA_CHOL_PERSISTENT", "0")
        ctx.ecrpa_batch(frags[:1], prec)   # back to the per-step mode for later tests


@pytest.mark.skip(reason="the persistent Cholesky launch was taken out of the library (it did not "
                         "complete on the GPU, twice); there is no switch left to test")
def test_persistent_cholesky_is_switched_off(ctx, monkeypatch):
    import brpa
    f = synth.make_fragment(3, 20, 100, 5)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, 8)
    monkeypatch.setenv("BRPA_CHOL_PERSISTENT", "1")
    try:
        with pytest.raises(brpa.BrpaError) as ei:
            ctx.ecrpa_batch([(f.R, f.eocc, f.evir, x, w)], "fp64")
        assert ei.value.code == brpa._lib.ERR_ARG
    finally:
        monkeypatch.setenv("BRPA_CHOL_PERSISTENT", "0")
        ctx.ecrpa_batch([(f.R, f.eocc, f.evir, x, w)], "fp64")
