"""The Fortran drop-in (brpa_mod, amdflang, 64-bit default INTEGER) called with
the reference's argument lists, checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from brpa import synth
from oracle import aomo, erpa, grid, mbpt3, sosex, t2, thc

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "beyond-rpa_amd", "brpa", "brpa_shim_check")


def test_fortran_shim_matches_oracle(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("brpa_shim_check was not built (amdflang missing at build time?)")
    o, v, n, nf = 5, 24, 60, 12
    f = synth.make_fragment(o, v, n, 99)
    x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.txt"
    with open(fin, "wb") as fh:
        np.array([o, v, n, nf, 0], dtype=np.int64).tofile(fh)
        # Fortran Rkai(NVecs, NVirt*NOcc) is C R[ai][P]
        for a in (f.R, f.eocc, f.evir, x, w):
            np.ascontiguousarray(a, np.float64).tofile(fh)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = open(fout).read().split()
    ec, efused, esplit, nt, ex = float(vals[0]), float(vals[1]), float(vals[2]), int(vals[3]), float(vals[4])
    A_f = np.array([float(s) for s in vals[5:5 + nt]])
    e_ref = erpa.ecrpa(f.R, f.gaps(), x, w)
    assert abs(ec - e_ref) <= 1e-8
    assert abs(efused - e_ref) <= 1e-8
    assert abs(esplit - e_ref) <= 1e-7
    Pis = [erpa.gpiug_mo(f.R, f.gaps(), xx) for xx in x]
    A_ref, V_ref = t2.t2_v5(f.R, f.gaps(), x, w, Pis, seed=0)
    assert nt == len(A_ref)
    assert np.max(np.abs(A_f - A_ref)) <= 1e-9 * np.max(np.abs(A_ref))
    assert abs(ex - sosex.sosex(f.R, V_ref, A_ref, o, v)) <= 1e-8
    # rpa_CC_Exchange_MBPT3_1 / _2 on the same amplitudes
    e31, e32 = float(vals[5 + nt]), float(vals[6 + nt])
    e31_ref = mbpt3.mbpt3_1(f.R, V_ref, A_ref, o, v, f.eocc, f.evir)
    e32_ref = mbpt3.mbpt3_2(f.R, V_ref, A_ref, o, v)
    assert abs(e31 - e31_ref) <= 1e-8 * abs(e31_ref)
    assert abs(e32 - e32_ref) <= 1e-8 * abs(e32_ref)
    # rpa_THC_CC_T2: EIG/MAXEIG rule (1e-3), smooth step 0.1, supersystem
    k = 7 + nt
    nth, common, esum = int(vals[k]), float(vals[k + 1]), float(vals[k + 2])
    A_th = np.array([float(s) for s in vals[k + 3:k + 3 + nth]])
    vnorm, t11 = float(vals[k + 3 + nth]), float(vals[k + 4 + nth])
    A_c, V_c, K, common_ref, _ = t2.t2_cutoff(A_ref, V_ref, t2.T2_CUTOFF_EIG_DIV_MAXEIG, 1e-3, o,
                                              smooth_step_on=True, steepness=0.1)
    assert nth == len(A_c)
    assert abs(common - common_ref) <= 1e-9 * abs(common_ref)
    assert np.max(np.abs(A_th - A_c)) <= 1e-9 * np.max(np.abs(A_c))
    assert abs(esum - t2.eig_rpa(A_c, V_c, f.R).sum()) <= 1e-9
    assert abs(vnorm - 1.0) <= 1e-9
    assert abs(t11 - np.sum(A_c * V_c[0, :] ** 2)) <= 1e-9 * np.max(np.abs(A_c))


def test_fortran_thc_piu_matches_oracle(tmp_path):
    """rpa_THC_PiU through the Fortran module (closed shell, FullRkai)."""
    if not os.path.exists(EXE):
        pytest.fail("brpa_shim_check was not built (amdflang missing at build time?)")
    o, v, n, g, nf = 3, 17, 40, 90, 5
    f = synth.make_fragment(2, 3, 4, 1)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.txt"
    with open(fin, "wb") as fh:
        np.array([f.nocc, f.nvir, f.naux, 4, 0], dtype=np.int64).tofile(fh)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, 4)
        for a in (f.R, f.eocc, f.evir, x, w):
            np.ascontiguousarray(a, np.float64).tofile(fh)
    t = synth.make_thc_fragment(o, v, n, g, 5)
    freqs = np.geomspace(0.05, 10.0, nf)
    tin, tout = tmp_path / "thc.bin", tmp_path / "thc.txt"
    with open(tin, "wb") as fh:
        np.array([o, v, n, g, nf], dtype=np.int64).tofile(fh)
        for a in (t.Xga, t.Xgi, t.Zgk, t.eocc, t.evir, freqs):
            np.ascontiguousarray(a, np.float64).tofile(fh)
    # rpa_CRG: shells s, p, d, f-like, two AO-pair subsets
    sizes = [1, 3, 5, 1, 3, 7, 1, 3]
    nao = sum(sizes)
    loc = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = aomo.pair_table([(A, B) for A in range(len(sizes)) for B in range(A + 1)], loc, sizes)
    rng = np.random.default_rng(11)
    nv, no_, nvr = 37, 3, 11
    RG = rng.standard_normal((len(pairs), nv))
    Co = rng.standard_normal((no_, nao))
    Cv = rng.standard_normal((nvr, nao))
    ain, aout = tmp_path / "aomo.bin", tmp_path / "aomo.txt"
    with open(ain, "wb") as fh:
        np.array([nao, len(sizes), nv, no_, nvr] + sizes, dtype=np.int64).tofile(fh)
        for a in (RG, Co, Cv):
            np.ascontiguousarray(a, np.float64).tofile(fh)
    r = subprocess.run([EXE, str(fin), str(fout), str(tin), str(tout), str(ain), str(aout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = np.array([float(s) for s in open(tout).read().split()])
    Pi = vals[:nf * n * n].reshape(nf, n, n)
    R = vals[nf * n * n:].reshape(o * v, n)
    Pis_ref, R_ref = thc.thc_piu(t.Xga.T, t.Xgi.T, t.Zgk.T, t.eocc, t.evir, freqs)
    assert np.max(np.abs(R - R_ref)) <= 1e-13 * np.max(np.abs(R_ref))
    assert np.max(np.abs(Pi - Pis_ref)) <= 1e-12 * np.max(np.abs(Pis_ref))
    CR = np.array([float(x) for x in open(aout).read().split()]).reshape(no_ * nvr, nv)
    ref = aomo.crg(RG, pairs, Co.T, Cv.T)
    assert np.max(np.abs(CR - ref)) <= 1e-13 * np.max(np.abs(ref))
