"""Writes tests/golden/golden.json: small known-answer cases of the E_RPA and
exchange path and of the rows around it (THC R and Pi, the THC amplitude
cutoff, MBPT3-1/2, THC SOSEX+2g, AO->MO, the Pi basis), computed by the numpy
oracle (the reference holds no golden vectors for this path and cannot be
compiled here).

Inputs are regenerated from (shape, seed) by brpa.synth; the SHA-256 of the
generated R bytes is stored so a change of the generator is caught.  Run:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "beyond-rpa_amd"))

import numpy as np  # noqa: E402

from brpa import synth  # noqa: E402
from oracle import aomo, basis, erpa, grid, mbpt3, sosex, t2, thc  # noqa: E402

# (name, nocc, nvir, naux, nfreqs, seed); the first is the CPU-runnable
# reference case o=8, v=72, NAux=400, 16 frequencies
ERPA_CASES = [("config1", 8, 72, 400, 16, 20261202), ("small_a", 3, 11, 70, 8, 5),
              ("n_gt_ov", 2, 9, 40, 8, 6), ("tiny", 2, 3, 5, 8, 7)]
# T2 + SOSEX: full-space cases (NG >= ov) and one truncated case
T2_CASES = [("t2_tiny", 2, 4, 10, 8, 11, None), ("t2_small", 3, 8, 30, 10, 12, None),
            ("t2_trunc", 4, 10, 45, 8, 13, 25)]


# THC factors (name, nocc, nvir, naux, ngrid, nfreqs, seed)
THC_CASES = [("thc_small", 3, 11, 40, 90, 8, 21), ("thc_ragged", 5, 27, 70, 133, 6, 22)]
# amplitude cutoff + MBPT3 (name, nocc, nvir, naux, nfreqs, seed, cutoff type, thresh, smooth)
CUTOFF_CASES = [("cut_eig", 3, 10, 30, 8, 31, 0, 1e-3, False),
                ("cut_maxeig_smooth", 4, 12, 40, 8, 32, 1, 1e-2, True),
                ("cut_sum_rejected", 3, 14, 36, 8, 33, 3, 1e-3, False)]
# THC SOSEX + 2g (name, nocc, nvir, naux, ngrid, nfreqs, seed)
THC_1B2G_CASES = [("c1b2g_small", 3, 9, 24, 70, 8, 41), ("c1b2g_ragged", 4, 13, 30, 101, 8, 42)]
# AO -> MO (name, shell sizes, nocc, nvir, nvecs, seed)
AOMO_CASES = [("aomo_spd", [1, 3, 5, 1, 3], 3, 7, 20, 51),
              ("aomo_spdf", [1, 1, 3, 3, 5, 7, 1, 3], 4, 11, 37, 52)]
# Pi basis (name, nocc, nvir, naux, nfreqs, seed, target as a fraction of Tr Pi)
RPAG_CASES = [("rpag_small", 4, 20, 60, 8, 61, 1e-3), ("rpag_mid", 6, 30, 120, 8, 62, 1e-4)]


def r_hash(R):
    return hashlib.sha256(np.ascontiguousarray(R, np.float64).tobytes()).hexdigest()


def main():
    out = {"erpa": [], "t2_sosex": []}
    for name, o, v, n, nf, seed in ERPA_CASES:
        f = synth.make_fragment(o, v, n, seed)
        x, w, info = grid.freq_grid(f.eocc, f.evir, nf)
        e, eu = erpa.ecrpa(f.R, f.gaps(), x, w, return_terms=True)
        out["erpa"].append({"name": name, "nocc": o, "nvir": v, "naux": n, "nfreqs": nf,
                            "seed": seed, "R_sha256": r_hash(f.R), "freqs": list(x),
                            "weights": list(w), "alpha": info["alpha"], "ecrpa": e,
                            "eu": [float(t) for t in eu]})
    for name, o, v, n, nf, seed, ng in T2_CASES:
        f = synth.make_fragment(o, v, n, seed)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
        g = f.gaps()
        Pis = [erpa.gpiug_mo(f.R, g, xx) for xx in x]
        A, V = t2.t2_v5(f.R, g, x, w, Pis, nguess=ng, seed=seed)[:2]
        ex = sosex.sosex(f.R, V, A, o, v)
        ex_s2 = sosex.sosex(f.R, V, A, o, v, s2_operator=True)
        Tm = V @ np.diag(A) @ V.T
        out["t2_sosex"].append({"name": name, "nocc": o, "nvir": v, "naux": n, "nfreqs": nf,
                                "seed": seed, "nguess": ng, "R_sha256": r_hash(f.R),
                                "A": list(A), "T2_frobenius": float(np.linalg.norm(Tm)),
                                "T2_trace": float(np.trace(Tm)), "sosex": ex, "sosex_s2": ex_s2})
    # THC route: R and Pi(u) from THC factors, E_RPA of that Pi
    out["thc"] = []
    for name, o, v, n, g, nf, seed in THC_CASES:
        f = synth.make_thc_fragment(o, v, n, g, seed)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
        Pis, R = thc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, x)
        out["thc"].append({"name": name, "nocc": o, "nvir": v, "naux": n, "ngrid": g,
                           "nfreqs": nf, "seed": seed, "Z_sha256": r_hash(f.Zgk),
                           "R_frobenius": float(np.linalg.norm(R)), "R_sum": float(R.sum()),
                           "pi_trace": [float(np.trace(P)) for P in Pis],
                           "ecrpa": erpa.ecrpa_analytic(Pis, x, w)})
    # amplitude cutoff of the THC route, MBPT3-1/2 on MO-route amplitudes
    out["t2_cutoff_mbpt3"] = []
    for name, o, v, n, nf, seed, ctype, thresh, smooth in CUTOFF_CASES:
        f = synth.make_fragment(o, v, n, seed)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
        g = f.gaps()
        Pis = [erpa.gpiug_mo(f.R, g, xx) for xx in x]
        A, V = t2.t2_v5(f.R, g, x, w, Pis, seed=seed)
        Ac, Vc, K, _, nsc = t2.t2_cutoff(A, V, ctype, thresh, o, smooth_step_on=smooth)
        out["t2_cutoff_mbpt3"].append({
            "name": name, "nocc": o, "nvir": v, "naux": n, "nfreqs": nf, "seed": seed,
            "R_sha256": r_hash(f.R), "cutoff_type": ctype, "thresh": thresh, "smooth": smooth,
            "K": K, "nvecs_kept": len(Ac), "n_scaled": nsc, "A_kept": list(Ac),
            "eig_rpa_sum": float(t2.eig_rpa(Ac, Vc, f.R).sum()),
            "mbpt3_1": mbpt3.mbpt3_1(f.R, V, A, o, v, f.eocc, f.evir),
            "mbpt3_2": mbpt3.mbpt3_2(f.R, V, A, o, v)})
    # THC SOSEX + 2g on oracle amplitudes of the THC-built R
    out["thc_1b2g"] = []
    for name, o, v, n, g, nf, seed in THC_1B2G_CASES:
        f = synth.make_thc_fragment(o, v, n, g, seed)
        R = thc.thc_rkai(f.Xga.T, f.Xgi.T, f.Zgk.T)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
        Pis = [erpa.gpiug_mo(R, f.gaps(), xx) for xx in x]
        A, V = t2.t2_v5(R, f.gaps(), x, w, Pis, seed=seed)
        U = V.reshape(o, v, len(A)).transpose(1, 0, 2)
        Zgh = f.Zgk.T @ f.Zgk
        e1, e2, _, _, _ = thc.thc_1b2g(Zgh, f.Xgi.T, f.Xga.T, U, A)
        out["thc_1b2g"].append({"name": name, "nocc": o, "nvir": v, "naux": n, "ngrid": g,
                                "nfreqs": nf, "seed": seed, "Z_sha256": r_hash(f.Zgk),
                                "nvecs_t2": len(A), "ec1b": e1, "ec2g": e2})
    # AO -> MO transform: shells, random coefficients and Cholesky vectors
    out["aomo"] = []
    for name, sizes, o, v, nvecs, seed in AOMO_CASES:
        loc = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        pairs = aomo.pair_table([(A_, B_) for A_ in range(len(sizes)) for B_ in range(A_ + 1)],
                                loc, sizes)
        nao = int(sum(sizes))
        rng = np.random.default_rng(seed)
        RG = rng.standard_normal((len(pairs), nvecs))
        Co = rng.standard_normal((o, nao))
        Cv = rng.standard_normal((v, nao))
        Rm = aomo.crg(RG, pairs, Co.T, Cv.T)
        wts = np.cos(np.arange(Rm.size)).reshape(Rm.shape)
        out["aomo"].append({"name": name, "shell_sizes": sizes, "nocc": o, "nvir": v,
                            "nvecs": nvecs, "seed": seed, "RG_sha256": r_hash(RG),
                            "CRG_frobenius": float(np.linalg.norm(Rm)),
                            "CRG_weighted_sum": float(np.sum(Rm * wts))})
    # Pi basis (rpa_G), eigen basis of Pi at the lowest frequency
    out["rpa_g"] = []
    for name, o, v, n, nf, seed, frac in RPAG_CASES:
        f = synth.make_fragment(o, v, n, seed)
        x, w, _ = grid.freq_grid(f.eocc, f.evir, nf)
        RWR = erpa.gpiug_mo(f.R, f.gaps(), x[0])
        target = frac * float(np.trace(RWR))
        _, R2, nv, err = basis.rpa_g(RWR, basis.BASIS_EIGEN, target)
        out["rpa_g"].append({"name": name, "nocc": o, "nvir": v, "naux": n, "nfreqs": nf,
                             "seed": seed, "R_sha256": r_hash(f.R), "target": target,
                             "nvecs": nv, "abs_error": err,
                             "eigenvalues_top": [float(t) for t in np.diag(R2)[:8]]})
    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
