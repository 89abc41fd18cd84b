"""Writes tests/golden/golden.json: small known-answer cases of the E_RPA and
exchange path, computed by the numpy oracle (the reference holds no golden
vectors for this path and cannot be compiled here).

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
from oracle import erpa, grid, sosex, t2  # noqa: E402

# (name, nocc, nvir, naux, nfreqs, seed); the first is the CPU-runnable
# reference case o=8, v=72, NAux=400, 16 frequencies
ERPA_CASES = [("config1", 8, 72, 400, 16, 20261202), ("small_a", 3, 11, 70, 8, 5),
              ("n_gt_ov", 2, 9, 40, 8, 6), ("tiny", 2, 3, 5, 8, 7)]
# T2 + SOSEX: full-space cases (NG >= ov) and one truncated case
T2_CASES = [("t2_tiny", 2, 4, 10, 8, 11, None), ("t2_small", 3, 8, 30, 10, 12, None),
            ("t2_trunc", 4, 10, 45, 8, 13, 25)]


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
    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
