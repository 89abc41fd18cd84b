"""GPU parity of the THC route's amplitudes (rpa_THC_CC_T2: small-eigenvalue
cutoff, EigRPA, V in the (ov, NT2) layout) against the CPU oracle.

Thresholds are placed between two retained eigenvalues (geometric mean), so
the kept count is decided far from rounding.  Tolerances as for the amplitude
path: 1e-9 relative on A, T2 and EigRPA (FP64), 1e-8 Ha on E_x.
"""
import numpy as np
import pytest

import brpa
from brpa import synth
from oracle import erpa, grid, sosex, t2

pytestmark = pytest.mark.gpu


def _setup(o, v, n, nf, seed):
    f = synth.make_fragment(o, v, n, seed)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    return f, x, w, Pis


def _between(a, b):
    return float(np.sqrt(abs(a) * abs(b)))


def _check(ctx, f, x, w, A_ref, V_ref, o, v, ctype, thresh, **kw):
    Ac, Vc, K, common_ref, nsc_ref = t2.t2_cutoff(A_ref, V_ref, ctype, thresh, o, **kw)
    gkw = dict(kw)
    smooth = gkw.pop("smooth_step_on", False)
    A, Vt, eig, common, nsc = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, seed=77,
                                         t2_cutoff_thresh=thresh, t2_cutoff_type=ctype,
                                         smooth_step=smooth, **gkw)
    print("type", ctype, "NT2", len(A_ref), "->", len(A), len(Ac), "K", common, common_ref)
    assert len(A) == len(Ac) and nsc == nsc_ref
    assert Vt.shape == (len(A), o * v)
    assert abs(common - common_ref) <= 1e-9 * abs(common_ref)
    assert np.max(np.abs(A - Ac)) <= 1e-9 * np.max(np.abs(Ac))
    V = np.ascontiguousarray(Vt.T)                  # V(ov, NT2) -> V[ai][mu]
    T_ref = (Vc * Ac) @ Vc.T
    assert np.max(np.abs((V * A) @ V.T - T_ref)) <= 1e-9 * np.max(np.abs(T_ref))
    e_ref = t2.eig_rpa(Ac, Vc, f.R)
    assert np.max(np.abs(eig - e_ref)) <= 1e-9 * np.max(np.abs(e_ref))
    return A, V


@pytest.mark.parametrize("o,v,n,nf", [(4, 25, 40, 12), (6, 40, 90, 16)])
def test_thc_t2_cutoff_rules_match_oracle(ctx, o, v, n, nf):
    f, x, w, Pis = _setup(o, v, n, nf, 31)
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=77)
    j = len(A_ref) // 2
    kmid = _between(A_ref[j - 1], A_ref[j])
    _check(ctx, f, x, w, A_ref, V_ref, o, v, t2.T2_CUTOFF_EIG, kmid)
    _check(ctx, f, x, w, A_ref, V_ref, o, v, t2.T2_CUTOFF_EIG_DIV_MAXEIG,
           kmid / np.max(np.abs(A_ref)))
    _check(ctx, f, x, w, A_ref, V_ref, o, v, t2.T2_CUTOFF_EIG_DIV_NELECTRON, kmid * 2 * o)
    tails = np.abs(np.cumsum(A_ref[::-1])[::-1])    # |A(i) + ... + A(n)|, 0-based i
    _check(ctx, f, x, w, A_ref, V_ref, o, v, t2.T2_CUTOFF_SUM_REJECTED,
           _between(tails[j], tails[j + 1]))
    # smooth step: the interval (K/10, K) holds some eigenvalues
    _check(ctx, f, x, w, A_ref, V_ref, o, v, t2.T2_CUTOFF_EIG, kmid, smooth_step_on=True,
           steepness=0.1)


def test_thc_t2_subsystem_threshold_and_sosex(ctx):
    """The supersystem's K applied to a subsystem; SOSEX of the kept amplitudes."""
    o, v, n, nf = 5, 30, 60, 12
    f, x, w, Pis = _setup(o, v, n, nf, 8)
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=77)
    common = _between(A_ref[len(A_ref) // 3], A_ref[len(A_ref) // 3 + 1])
    Ac, Vc, _, _, _ = t2.t2_cutoff(A_ref, V_ref, t2.T2_CUTOFF_EIG_DIV_MAXEIG, 0.5, o,
                                   common_thresh=common)
    A, Vt, eig, common_out, _ = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, seed=77,
                                           t2_cutoff_thresh=0.5, t2_cutoff_type="eig/maxeig",
                                           common_thresh=common)
    assert common_out == common and len(A) == len(Ac)
    V = np.ascontiguousarray(Vt.T)
    ex = ctx.sosex(V, A, f.R, o, v)
    assert abs(ex - sosex.sosex(f.R, Vc, Ac, o, v)) <= 1e-8


def test_thc_t2_split_i8(ctx):
    """Split-int8 amplitude contractions followed by the cutoff: same kept set,
    T2 and EigRPA within 1e-7 relative of the FP64 oracle."""
    o, v, n, nf = 6, 40, 90, 12
    f, x, w, Pis = _setup(o, v, n, nf, 31)
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=77)
    kmid = _between(A_ref[9], A_ref[10])
    Ac, Vc, _, _, _ = t2.t2_cutoff(A_ref, V_ref, t2.T2_CUTOFF_EIG, kmid, o)
    A, Vt, eig, _, _ = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, seed=77, t2_cutoff_thresh=kmid,
                                  precision="split_i8")
    assert len(A) == len(Ac) == 10
    V = np.ascontiguousarray(Vt.T)
    T_ref = (Vc * Ac) @ Vc.T
    assert np.max(np.abs((V * A) @ V.T - T_ref)) <= 1e-7 * np.max(np.abs(T_ref))
    e_ref = t2.eig_rpa(Ac, Vc, f.R)
    assert np.max(np.abs(eig - e_ref)) <= 1e-7 * np.max(np.abs(e_ref))


def test_thc_t2_eig_rpa_full_space_equals_erpa(ctx):
    """No vector dropped (K = 0 from SUM-REJECTED with a tiny threshold), NG = ov:
    sum EigRPA = E_RPA of the same grid."""
    f, x, w, Pis = _setup(2, 3, 4, 24, 7)
    A, Vt, eig, common, _ = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, nguess=6, cutoff=0.0, seed=3,
                                       t2_cutoff_thresh=1e-300, t2_cutoff_type="sum-rejected")
    assert common == 0.0 and len(A) == 6
    assert abs(eig.sum() - erpa.ecrpa(f.R, f.gaps(), x, w)) < 1e-10


def test_thc_t2_empty_kept_set_is_an_error(ctx):
    f, x, w, _ = _setup(3, 10, 12, 10, 2)
    with pytest.raises(brpa.BrpaError):
        ctx.thc_t2(f.R, f.eocc, f.evir, x, w, t2_cutoff_thresh=10.0)
    with pytest.raises(brpa.BrpaError):
        ctx.thc_t2(f.R, f.eocc, f.evir, x, w, t2_cutoff_type=9)
    # the context stays usable
    A, _, _, _, _ = ctx.thc_t2(f.R, f.eocc, f.evir, x, w, t2_cutoff_thresh=1e-12)
    assert len(A) > 0
