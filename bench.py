"""Throughput of the batched correlation-energy hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision split_i8|split_bf16|fp64]

Workload: 256 uneven fragments, o ~ U{8..40}, v ~ U{70..350},
N = round(4.5 (o+v)), 16 frequencies each, seeded synthetic inputs
(brpa.synth), R already resident in HBM.  One step = E_RPA of all 256
fragments in batched launches (Pi(u) build, exact diagonal/trace, blocked FP64
Cholesky log-det, quadrature; replayed as a hipGraph per workspace chunk, one
eager pass supplies the per-stage kernel times), then the exchange correction
of all 256 fragments: the direct-ring amplitudes (T2, guess count ceil(3 N))
and the SOSEX contraction per fragment in the explicit ov space
(brpa_exchange_batch, four fragments in flight on worker threads, each on a
quarter of the CUs).  The headline form (--precision,
split-int8 by default) runs W warm-up steps and exactly K timed steps; the
other forms (split-BF16 and FP64, their exchange in FP64) are timed over one
step each and reported under `forms`.  The frequency grids (host code,
independent of R) are built once before timing.

Other workloads (--workload): single, adjoint, sosex, and thc (R from
tensor-hypercontraction factors and Pi(u) of all frequencies, one fragment),
mbpt3 (the MBPT3-1 and MBPT3-2 exchange corrections of the large fragment),
aomo (the AO->MO transform of the Cholesky vectors, one fragment), thc1b2g
(the THC SOSEX + 2g cumulant terms, one fragment).

Prints one JSON line.  `roofline` is measured for the dominant kernel of the job,
the FP64 MFMA GEMM of the explicit-space exchange (k_gemm_f64), with HIP events
inside the library around every launch of a one-worker exchange pass over
fragments 9..16;
`roofline_erpa` is the E_RPA product kernel's over the timed steps; `cpu_baseline` times the
CPU oracle (numpy float64 + threaded OpenBLAS: E_RPA, T2 v5 and SOSEX) on the
first fragments of the batch for about --cpu-seconds and checks the GPU
energies and exchange energies against it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "beyond-rpa_amd"))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X dense FP64 matrix
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense BF16 matrix
PEAK_I8_TOPS = 5000.0        # MI355X dense INT8 matrix
SPLIT_PRODUCTS = 10          # slice products per FP64-equivalent product (4 slices, level <= 3)
PRECISIONS = ("split_i8", "split_bf16", "fp64")
DTYPES = {"split_i8": "i8 (4-slice split, 10 products, exact i32 acc) + f64 log-det",
          "split_bf16": "bf16 (4-slice split, 10 products, fp32 acc) + f64 log-det",
          "fp64": "f64"}
# dominant Pi kernel and the peak its algorithmic FP64-equivalent flops are priced at
PI_KERNEL = {"split_i8": ("k_syrk_i8", PEAK_I8_TOPS / SPLIT_PRODUCTS),
             "split_bf16": ("k_syrk_bf16", PEAK_BF16_TFLOPS / SPLIT_PRODUCTS),
             "fp64": ("k_syrk_f64", PEAK_FP64_TFLOPS)}


def pmc_traffic(kernel, profile="erpa_batch256_split_i8_kernels.md"):
    """HBM bytes per launch (read + written) of `kernel` from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == kernel:
            try:
                return round((float(cells[4]) + float(cells[5])) * 1e9)
            except ValueError:
                return None
    return None


def syrk_flops(o, v, n, nf):
    """Algorithmic Pi(u) flops of one fragment: F * ov * N (N+1) (SYRK, FMA = 2)."""
    return nf * o * v * n * (n + 1)


def chol_flops(n, nf):
    return nf * n ** 3 / 3.0


def bench_sosex(a):
    """SOSEX contraction of one large fragment (o=64, v=600, NAux=3200, 16
    frequencies, amplitudes with NGuess = 3 NAux built once in FP64): FP64
    (rocBLAS pair GEMMs) vs the fused split-int8 kernel, |dE| between them."""
    import math

    import numpy as np
    import torch

    import brpa
    from brpa import synth
    _one_gpu(a)
    o, v, n = 64, 600, 3200
    f = synth.make_fragment(o, v, n, synth.BASE_SEED)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, 16)
    Rd = torch.from_numpy(f.R).cuda()
    ctx = brpa.Context(0)
    A, V = ctx.t2(Rd, f.eocc, f.evir, x, w, nguess=math.ceil(3.0 * n), seed=1, device=True)
    nt = A.numel()
    res = {}
    for prec in ("split_i8", "fp64"):
        for _ in range(a.warmup):
            ctx.sosex(V, A, Rd, o, v, precision=prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        kms = 0.0
        for _ in range(a.steps):
            e = ctx.sosex(V, A, Rd, o, v, precision=prec)
            kms += ctx.last_kernel_ms()
        torch.cuda.synchronize()
        res[prec] = (e, (time.perf_counter() - t) / a.steps, kms / a.steps)
        print("# %s: E_x %.10f  %.1f ms/step (kernels %.1f ms)" % (prec, e, res[prec][1] * 1e3,
                                                                  res[prec][2]), file=sys.stderr)
    flops = 2.0 * 2 * (o * (o + 1) // 2) * n * nt * v   # two W GEMMs per pair (i >= j)
    e8, dt8, k8 = res["split_i8"]
    e64, dt64, k64 = res["fp64"]
    out = {"metric": "E_SOSEX contraction time, split-int8 vs FP64 at |dE|<=1e-7 Ha",
           "value": round(dt8 * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(dt8 * 1e3, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None,
           "dtype": "i8 (5-slice split, 15 products, exact i32 acc) + f64 tile sums",
           "data": "synthetic (seeded PCG64, brpa.synth)",
           "config": {"workload": "SOSEX of one fragment o=64, v=600, NAux=3200, 16 "
                                  "frequencies, NT2=%d" % nt},
           "fp64_ms": round(dt64 * 1e3, 3), "speedup_vs_fp64": round(dt64 / dt8, 3),
           "abs_dE_vs_fp64_Ha": abs(e8 - e64), "E_x_fp64": e64,
           "roofline": {"bound": "mfma", "achieved": round(flops / (k8 * 1e-3) / 1e12, 3),
                        "peak": round(PEAK_I8_TOPS / 15, 3), "unit": "TFLOP/s",
                        "frac": round(flops / (k8 * 1e-3) / 1e12 / (PEAK_I8_TOPS / 15), 4),
                        "traffic": None, "kernel": "k_sosex_i8 (+ slicing)",
                        "fp64_achieved": round(flops / (k64 * 1e-3) / 1e12, 3)}}
    print(json.dumps(out))


def _one_gpu(a):
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)


def _cpu_threads():
    from threadpoolctl import threadpool_info
    return max([p.get("num_threads", 1) for p in threadpool_info()] or [1])


def _host_cpus():
    """Sockets x cores per socket (lscpu) and the CPUs this process may use."""
    info = {}
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True,
                                   timeout=10).stdout.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    sockets, cps = info.get("Socket(s)"), info.get("Core(s) per socket")
    return {"lscpu_sockets": int(sockets) if sockets and sockets.isdigit() else None,
            "lscpu_cores_per_socket": int(cps) if cps and cps.isdigit() else None,
            "model": info.get("Model name"),
            "cpus_available_to_process": len(os.sched_getaffinity(0))}


def bench_single(a):
    """E_RPA of one large fragment (o=36, v=300, NAux=1800, 24 frequencies) in
    every form on one GPU, each energy against the FP64 CPU oracle; the oracle's
    own time on the host cores is the CPU baseline."""
    import torch

    import brpa
    from brpa import synth
    from oracle import erpa as oerpa
    _one_gpu(a)
    o, v, n, nf = 36, 300, 1800, 24
    f = synth.make_fragment(o, v, n, synth.BASE_SEED)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, nf)
    Rd = torch.from_numpy(f.R).cuda()
    torch.cuda.synchronize()
    ctx = brpa.Context(0)
    res = {}
    for prec in (a.precision,) + tuple(p for p in PRECISIONS if p != a.precision):
        ctx.set_graphs(False)
        e_eager = ctx.ecrpa(Rd, f.eocc, f.evir, x, w, prec)[0]
        stage = {k: s[0] for k, s in ctx.last_stage_ms().items()}
        ctx.set_graphs(bool(a.graphs))
        for _ in range(a.warmup):
            ctx.ecrpa(Rd, f.eocc, f.evir, x, w, prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.steps):
            e = ctx.ecrpa(Rd, f.eocc, f.evir, x, w, prec)[0]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / a.steps
        ctx.set_graphs(False)
        if e != e_eager:
            raise RuntimeError("%s: replayed energy differs from the eager pass" % prec)
        res[prec] = {"E": e, "ms": dt * 1e3, "stage_ms": stage}
        print("# %s: E %.10f  %.2f ms/step stages %s" % (prec, e, dt * 1e3, stage),
              file=sys.stderr, flush=True)
    t = time.perf_counter()
    eref = oerpa.ecrpa(f.R, f.gaps(), x, w)
    tc = time.perf_counter() - t
    for d in res.values():
        d["abs_dE_vs_cpu_oracle_Ha"] = abs(d["E"] - eref)
    main = res[a.precision]
    flops = syrk_flops(o, v, n, nf)
    kernel, peak = PI_KERNEL[a.precision]
    achieved = flops / (main["stage_ms"]["pi"] * 1e-3) / 1e12
    out = {"metric": "E_RPA time of one fragment o=36, v=300, NAux=1800, 24 frequencies, "
                     "split vs FP64 at |dE|<=1e-7 Ha",
           "value": round(main["ms"], 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(main["ms"], 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": DTYPES[a.precision],
           "data": "synthetic (seeded PCG64, brpa.synth)",
           "config": {"workload": "single fragment o=36, v=300, NAux=1800, 24 frequencies",
                      "precision": a.precision, "graphs": bool(a.graphs)},
           "E_cpu_oracle": eref,
           "speedup_vs_fp64": round(res["fp64"]["ms"] / main["ms"], 3),
           "forms": {p: {"ms_per_step": round(d["ms"], 3), "E": d["E"],
                         "abs_dE_vs_cpu_oracle_Ha": d["abs_dE_vs_cpu_oracle_Ha"],
                         "stage_ms": {k: round(s, 3) for k, s in d["stage_ms"].items()}}
                     for p, d in res.items()},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 3),
                        "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": None,
                        "kernel": kernel, "kernel_ms_per_step": round(main["stage_ms"]["pi"], 3)},
           "cpu_baseline": {"value": round(tc * 1e3, 1), "unit": "ms", "cores": _cpu_threads(),
                            "kind": "port",
                            "sample": "the same fragment, E_RPA by the numpy FP64 oracle "
                                      "(dgemm Pi + dgeqrf log-det), one evaluation"}}
    print(json.dumps(out))


def bench_adjoint(a):
    """dE_RPA/dR and dE_RPA/d eps of one fragment (o=48, v=400, NAux=2400, 16
    frequencies, Pi recomputed in the gradient call) on one GPU; checked against
    central finite differences on --fd-entries random R entries plus 6 orbital
    energies and against the FP64 closed form on 8 entries."""
    import numpy as np
    import torch

    import brpa
    from brpa import synth
    from oracle import adjoint
    _one_gpu(a)
    o, v, n, nf = 48, 400, 2400, 16
    f = synth.make_fragment(o, v, n, seed=20261202)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, nf)
    Rd = torch.from_numpy(f.R).cuda()
    torch.cuda.synchronize()
    ctx = brpa.Context(0)
    for _ in range(a.warmup):
        ctx.ecrpa_grad(Rd, f.eocc, f.evir, x, w)
    torch.cuda.synchronize()
    t = time.perf_counter()
    kms = 0.0
    for _ in range(a.steps):
        dR, deo, dev = ctx.ecrpa_grad(Rd, f.eocc, f.evir, x, w)
        kms += ctx.last_kernel_ms()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / a.steps
    kms /= a.steps
    ctx.ecrpa(Rd, f.eocc, f.evir, x, w, "fp64")
    t_e = ctx.last_kernel_ms()
    print("# adjoint: %.1f ms/step (kernels %.1f ms), energy alone %.1f ms" % (dt * 1e3, kms, t_e),
          file=sys.stderr, flush=True)
    dR = dR.cpu().numpy()
    # finite differences need (1 + Pi_k)^-1; the Pi_k come from the GPU build
    # (its parity against the oracle is tested separately), the inverses and
    # the differences from the CPU
    t = time.perf_counter()
    Pis = []
    for xx in x:
        P = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        ctx.gpiug_mo(P, Rd, f.eocc, f.evir, xx, 4.0)
        Pis.append(P.cpu().numpy())
    Ms = adjoint.inverses(f.R, f.eocc, f.evir, x, Pis=Pis)
    rng = np.random.default_rng(2024)
    ents = [(int(rng.integers(0, o * v)), int(rng.integers(0, n))) for _ in range(a.fd_entries)]
    fd = adjoint.central_fd_R(f.R, f.eocc, f.evir, x, w, ents, 1e-5, Ms=Ms)
    worst = max(abs(fd[e] - dR[e]) / abs(fd[e]) for e in ents)
    for which, idx, g in (("o", 0, deo), ("o", o // 2 - 1, deo), ("o", o - 1, deo),
                          ("v", 0, dev), ("v", v // 2 - 1, dev), ("v", v - 1, dev)):
        fv = adjoint.central_fd_eps(f.R, f.eocc, f.evir, x, w, which, idx, 1e-5, Ms=Ms)
        worst = max(worst, abs(fv - g[idx]) / abs(fv))
    ref = adjoint.ecrpa_grad_R_entries(f.R, f.eocc, f.evir, x, w, ents[:8], Ms=Ms)
    worst_cf = max(abs(dR[e] - g) / abs(g) for e, g in ref.items())
    tcheck = time.perf_counter() - t
    print("# FD worst relative %.2e, closed form worst relative %.2e (%.1f s)" %
          (worst, worst_cf, tcheck), file=sys.stderr, flush=True)
    # per frequency, algorithmic: Pi SYRK (ov N^2), Cholesky (N^3/3), inverse from
    # the factor (2 N^3/3), M R^T (2 ov N^2), dR accumulation (2 ov N)
    flops = nf * (o * v * n * (n + 1) + n ** 3 / 3.0 + 2.0 * n ** 3 / 3.0 + 2.0 * o * v * n * n)
    achieved = flops / (kms * 1e-3) / 1e12
    # CPU baseline: the oracle's gradient on a bounded sample of the frequencies
    thr = _cpu_threads()
    t = time.perf_counter()
    k = 0
    while k < nf and (k == 0 or time.perf_counter() - t < a.cpu_seconds):
        adjoint.ecrpa_grad(f.R, f.eocc, f.evir, x[k:k + 1], w[k:k + 1])
        k += 1
    tc = (time.perf_counter() - t) / k * nf
    out = {"metric": "dE_RPA/dR + dE_RPA/d eps time of one fragment o=48, v=400, NAux=2400, "
                     "16 frequencies (FD-checked to 1e-6 relative)",
           "value": round(dt * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded PCG64, brpa.synth)",
           "config": {"workload": "adjoint of one fragment o=48, v=400, NAux=2400, 16 "
                                  "frequencies, Pi recomputed in the gradient call"},
           "kernel_ms_per_step": round(kms, 3), "energy_only_kernel_ms": round(t_e, 3),
           "fd_check": {"entries_R": len(ents), "orbital_energies": 6, "h": 1e-5,
                        "worst_relative": worst, "closed_form_entries": len(ref),
                        "closed_form_worst_relative": worst_cf, "pass": bool(worst <= 1e-6)},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 3),
                        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                        "kernel": "whole adjoint (SYRK + batched Cholesky + triangular inverse + dgemm + accumulation)"},
           "cpu_baseline": {"value": round(tc * 1e3, 1), "unit": "ms", "cores": thr,
                            "kind": "port",
                            "sample": "numpy FP64 oracle gradient on %d of the %d frequencies, "
                                      "scaled to all %d" % (k, nf, nf)}}
    print(json.dumps(out))


def bench_thc(a):
    """THC route: Cholesky vectors R from the THC factors and Pi(u) of all
    frequencies (brpa_thc_piu) for one fragment o=36, v=300, NAux=1800 with a
    4032-point THC grid and 16 frequencies, factors resident in HBM.  The CPU
    baseline is the numpy oracle of the same job (R, then Pi of every frequency)."""
    import numpy as np
    import torch

    import brpa
    from brpa import synth
    from oracle import erpa as oerpa, thc as othc
    _one_gpu(a)
    o, v, n, g, nf = synth.THC_CONFIG
    f = synth.make_thc_fragment(o, v, n, g, synth.BASE_SEED)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, nf)
    d = lambda t: torch.from_numpy(t).cuda()  # noqa: E731
    Xa, Xi, Z = d(f.Xga), d(f.Xgi), d(f.Zgk)
    Pi = torch.empty((nf, n, n), dtype=torch.float64, device="cuda")
    R = torch.empty((o * v, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx = brpa.Context(0)
    run = lambda: ctx.thc_piu(Xa, Xi, Z, f.eocc, f.evir, x, out=Pi, want_R=True,  # noqa: E731
                              R_out=R)
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    st_r = st_pi = 0.0
    t = time.perf_counter()
    for _ in range(a.steps):
        run()
        st = ctx.last_thc_stage_ms()
        st_r += st["thc_r"][0]
        st_pi += st["thc_pi"][0]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / a.steps
    st_r /= a.steps
    st_pi /= a.steps
    e_gpu = ctx.ecrpa_analytic(Pi, x, w)[0]
    print("# THC: %.3f ms/step, R kernel %.3f ms, Pi SYRK %.3f ms" % (dt * 1e3, st_r, st_pi),
          file=sys.stderr, flush=True)
    tc = None
    dpi = de = None
    if a.cpu_seconds > 0:
        t = time.perf_counter()
        Pis_ref, R_ref = othc.thc_piu(f.Xga.T, f.Xgi.T, f.Zgk.T, f.eocc, f.evir, x)
        tc = time.perf_counter() - t
        dpi = float(np.max(np.abs(Pi.cpu().numpy() - Pis_ref)) / np.max(np.abs(Pis_ref)))
        de = abs(e_gpu - oerpa.ecrpa_analytic(Pis_ref, x, w))
        del Pis_ref, R_ref
    r_flops = 2.0 * o * v * n * g
    pi_flops = syrk_flops(o, v, n, nf)
    r_tf = r_flops / (st_r * 1e-3) / 1e12
    pi_tf = pi_flops / (st_pi * 1e-3) / 1e12
    dom = ("k_syrk_f64", pi_tf, st_pi) if st_pi >= st_r else ("k_thc_r", r_tf, st_r)
    oth = ("k_thc_r", r_tf, st_r) if dom[0] == "k_syrk_f64" else ("k_syrk_f64", pi_tf, st_pi)
    out = {"metric": "THC-route Pi(u) build (R from the THC factors, then Pi of all 16 "
                     "frequencies) of one fragment o=36, v=300, NAux=1800, NGridTHC=4032",
           "value": round(dt * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded PCG64, brpa.synth.make_thc_fragment)",
           "config": {"workload": "THC fragment o=36, v=300, NAux=1800, NGridTHC=4032, "
                                  "16 frequencies", "precision": "fp64"},
           "stage_ms": {"thc_r": round(st_r, 3), "thc_pi": round(st_pi, 3)},
           "E_RPA_from_Pi": e_gpu, "max_rel_dPi_vs_cpu_oracle": dpi,
           "abs_dE_vs_cpu_oracle_Ha": de,
           "roofline": {"bound": "mfma", "achieved": round(dom[1], 3), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(dom[1] / PEAK_FP64_TFLOPS, 4),
                        "traffic": None, "kernel": dom[0], "kernel_ms_per_step": round(dom[2], 3)},
           "roofline_other": {"kernel": oth[0], "achieved": round(oth[1], 3),
                              "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(oth[1] / PEAK_FP64_TFLOPS, 4),
                              "kernel_ms_per_step": round(oth[2], 3)},
           "algorithmic_flop": {"thc_r": r_flops, "thc_pi": pi_flops},
           "cpu_baseline": None if tc is None else {
               "value": round(tc * 1e3, 1), "unit": "ms", "cores": _cpu_threads(), "kind": "port",
               "sample": "the same fragment, R and Pi of all 16 frequencies by the numpy FP64 "
                         "oracle, one evaluation"}}
    print(json.dumps(out))


def bench_mbpt3(a):
    """Third-order exchange corrections MBPT3-1 and MBPT3-2 of the large SOSEX
    fragment (o=64, v=600, NAux=3200, 16 frequencies, amplitudes with NGuess =
    3 NAux built once in FP64 on the GPU).  CPU baseline: the oracle's loops
    on a bounded sample (a subset of occupied pairs / (c, k) pairs and a
    smaller Lyapunov solve), scaled to the whole job by operation counts."""
    import math

    import numpy as np
    import torch

    import brpa
    from brpa import synth
    from oracle import mbpt3 as omb
    _one_gpu(a)
    o, v, n = 64, 600, 3200
    f = synth.make_fragment(o, v, n, synth.BASE_SEED)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, 16)
    Rd = torch.from_numpy(f.R).cuda()
    ctx = brpa.Context(0)
    A, V = ctx.t2(Rd, f.eocc, f.evir, x, w, nguess=math.ceil(3.0 * n), seed=1, device=True)
    nt = A.numel()
    res = {}
    for name, fn in (("mbpt3_1_split_i8", lambda: ctx.mbpt3_1(V, A, Rd, f.eocc, f.evir,
                                                              precision="split_i8")),
                     ("mbpt3_1_fp64", lambda: ctx.mbpt3_1(V, A, Rd, f.eocc, f.evir)),
                     ("mbpt3_2_fp64", lambda: ctx.mbpt3_2(Rd, V, A, o, v))):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st_core = st_sx = 0.0
        iters = 0
        for _ in range(a.steps):
            e = fn()
            st = ctx.last_mbpt3_stage_ms()
            st_core += st["mbpt3"][0]
            st_sx += st["sosex"][0]
            iters = st["mbpt3"][1]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / a.steps
        res[name] = {"E": e, "ms": dt * 1e3, "core_ms": st_core / a.steps,
                     "contraction_ms": st_sx / a.steps, "sign_iterations": iters}
        print("# %s: E %.10f  %.1f ms/step (core %.1f ms, contraction %.1f ms, iters %d)"
              % (name, e, dt * 1e3, st_core / a.steps, st_sx / a.steps, iters), file=sys.stderr,
              flush=True)
    K = o * v
    r1 = res["mbpt3_1_split_i8"]
    # algorithmic FP64 flop of the MBPT3-1 core: W, U, D, then per sign step the
    # Newton step's getrf + getri (2 n^3) and two GEMMs (4 n^3); the final
    # Newton-Schulz steps (six GEMMs) are counted at the same 6 n^3
    core_flops = 2.0 * K * n * nt + 2.0 * n * nt * nt + 2.0 * K * nt * nt \
        + 6.0 * nt ** 3 * max(r1["sign_iterations"], 1)
    m2_flops = 2.0 * o * o * v * v * (nt + n)
    m2 = res["mbpt3_2_fp64"]
    m2_tf = m2_flops / (m2["core_ms"] * 1e-3) / 1e12
    core_tf = core_flops / (r1["core_ms"] * 1e-3) / 1e12
    # CPU baseline on a bounded sample (one thread pool of the host)
    cpu = None
    if a.cpu_seconds > 0:
        Vh = V.cpu().numpy()
        Ah = A.cpu().numpy()
        rng = np.random.default_rng(0)
        t = time.perf_counter()
        # MBPT3-2: the reference loop over (c, k) on 2 of the 600 virtual indices
        R3 = f.R
        V3 = Vh.reshape(o, v, nt)
        cs = rng.choice(v, 2, replace=False)
        tr = 0.0
        for c in cs:
            Vli = Ah[:, None] * V3[:, c, :].T
            for k in range(o):
                Wai = V3[k] @ Vli
                Wm = R3.T @ Wai.T.ravel()
                tr += float(np.sum(Wm * Wm))
        t_m2 = (time.perf_counter() - t) * v / len(cs)
        # MBPT3-1: TrXK pair loop on 2 of the o(o+1)/2 pairs, Lyapunov at nt/4 (n^3)
        t = time.perf_counter()
        Rr = f.R.reshape(o, v, n)
        npair = o * (o + 1) // 2
        for (i, j) in ((1, 0), (5, 2)):
            Wij = Rr[i].T @ V3[j]
            Wji = Rr[j].T @ V3[i]
            np.sum(Wji.T @ Wij)
        t_pairs = (time.perf_counter() - t) * npair / 2
        m = nt // 4
        B = rng.standard_normal((m, m)) / math.sqrt(m) + 4.0 * np.eye(m)
        S = rng.standard_normal((m, m))
        t = time.perf_counter()
        omb.lyapunov(B, S + S.T)
        t_lyap = (time.perf_counter() - t) * 64.0
        cpu = {"mbpt3_1_ms": round((t_pairs + t_lyap) * 1e3, 1), "mbpt3_2_ms": round(t_m2 * 1e3, 1)}
    total = r1["ms"] + m2["ms"]
    out = {"metric": "MBPT3-1 + MBPT3-2 exchange time of one fragment o=64, v=600, NAux=3200, "
                     "NT2=%d (split-int8 final contraction of MBPT3-1)" % nt,
           "value": round(total, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(total, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None,
           "dtype": "f64 (Lyapunov, MBPT3-2) + i8 split (MBPT3-1 contraction)",
           "data": "synthetic (seeded PCG64, brpa.synth)",
           "config": {"workload": "MBPT3 exchange of one fragment o=64, v=600, NAux=3200, 16 "
                                  "frequencies, NT2=%d" % nt},
           "forms": {k: {kk: (round(vv, 3) if isinstance(vv, float) and kk != "E" else vv)
                         for kk, vv in d.items()} for k, d in res.items()},
           "abs_dE_mbpt3_1_split_vs_fp64_Ha": abs(res["mbpt3_1_split_i8"]["E"] -
                                                  res["mbpt3_1_fp64"]["E"]),
           "roofline": {"bound": "mfma", "achieved": round(m2_tf, 3), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(m2_tf / PEAK_FP64_TFLOPS, 4),
                        "traffic": None, "kernel": "MBPT3-2 products (rocBLAS dgemm)",
                        "kernel_ms_per_step": round(m2["core_ms"], 3)},
           "roofline_mbpt3_1_core": {"achieved": round(core_tf, 3), "peak": PEAK_FP64_TFLOPS,
                                     "unit": "TFLOP/s",
                                     "frac": round(core_tf / PEAK_FP64_TFLOPS, 4),
                                     "ms": round(r1["core_ms"], 3)},
           "cpu_baseline": None if cpu is None else {
               "value": round(cpu["mbpt3_1_ms"] + cpu["mbpt3_2_ms"], 1), "unit": "ms",
               "cores": _cpu_threads(), "kind": "port", "parts": cpu,
               "sample": "numpy oracle loops: MBPT3-2 over 2 of 600 virtual indices (x300), "
                         "the MBPT3-1 pair loop over 2 of %d pairs (x%d/2), the Lyapunov "
                         "solve at NT2/4 (x64, n^3)" % (o * (o + 1) // 2, o * (o + 1) // 2)}}
    print(json.dumps(out))


def bench_aomo(a):
    """AO -> MO transform of the Cholesky vectors (brpa_ao_mo: unpack to full AO-pair
    matrices, two quarter transforms) of one fragment:
    o=36, v=300 active orbitals over NAO=336 basis functions in s/p/d/f-like
    shells, NVecs=1800, all NAO(NAO+1)/2 packed AO pairs in one subset; inputs
    resident in HBM.  CPU baseline: the numpy oracle (C matrix, then one GEMM)
    on 2 of the 36 occupied orbitals, scaled by 18."""
    import numpy as np
    import torch

    import brpa
    from oracle import aomo
    _one_gpu(a)
    o, v, nvecs = 36, 300, 1800
    sizes = ([1, 1, 3, 3, 5, 7] * 17)[:-1]
    sizes = sizes + [336 - sum(sizes)] if sum(sizes) < 336 else sizes
    nao = sum(sizes)
    loc = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = aomo.pair_table([(A, B) for A in range(len(sizes)) for B in range(A + 1)], loc, sizes)
    rng = np.random.default_rng(7)
    RG = rng.standard_normal((len(pairs), nvecs)) * 0.01
    Co = rng.standard_normal((o, nao)) / np.sqrt(nao)
    Cv = rng.standard_normal((v, nao)) / np.sqrt(nao)
    d = lambda t: torch.from_numpy(t).cuda()  # noqa: E731
    RGd, Cod, Cvd = d(RG), d(Co), d(Cv)
    out = torch.empty((o * v, nvecs), dtype=torch.float64, device="cuda")
    ctx = brpa.Context(0)
    for _ in range(a.warmup):
        ctx.ao_mo(RGd, pairs, Cod, Cvd, out=out)
    torch.cuda.synchronize()
    kms = 0.0
    t = time.perf_counter()
    for _ in range(a.steps):
        ctx.ao_mo(RGd, pairs, Cod, Cvd, out=out)
        kms += ctx.last_kernel_ms()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / a.steps
    kms /= a.steps
    # algorithmic work of the two quarter transforms (the reference's packed form
    # costs 2 ov Npq NVecs, about NAO v / (2 (NAO + v)) times more)
    flops = 2.0 * nvecs * nao * o * (nao + v)
    ref_flops = 2.0 * o * v * len(pairs) * nvecs
    tf = flops / (kms * 1e-3) / 1e12
    cpu = None
    if a.cpu_seconds > 0:
        t = time.perf_counter()
        ref = aomo.crg(RG, pairs, Co[:2].T, Cv.T)
        tc = (time.perf_counter() - t) * (o / 2)
        got = out[: 2 * v].cpu().numpy()
        dmax = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        cpu = (tc, dmax)
    print("# AO->MO: %.3f ms/step, kernel %.3f ms, %.1f TF/s" % (dt * 1e3, kms, tf),
          file=sys.stderr, flush=True)
    res = {"metric": "AO->MO transform of the Cholesky vectors, one fragment o=36, v=300, "
                     "NAO=%d, NVecs=1800, %d packed AO pairs" % (nao, len(pairs)),
           "value": round(dt * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded PCG64 coefficients and Cholesky vectors)",
           "config": {"workload": "AO->MO transform o=36, v=300, NAO=%d, NVecs=1800" % nao},
           "max_rel_dCRG_vs_cpu_oracle_2_occupied": None if cpu is None else cpu[1],
           "algorithmic_flop": flops, "reference_form_flop": ref_flops,
           "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4),
                        "traffic": None,
                        "kernel": "brpa_ao_mo (k_unpack_pairs + 2 strided-batched dgemm)",
                        "kernel_ms_per_step": round(kms, 3)},
           "cpu_baseline": None if cpu is None else {
               "value": round(cpu[0] * 1e3, 1), "unit": "ms", "cores": _cpu_threads(),
               "kind": "port", "sample": "numpy oracle (C matrix + GEMM) on 2 of 36 occupied "
                                         "orbitals, x18"}}
    print(json.dumps(res))


def bench_thc1b2g(a):
    """THC SOSEX + 2g cumulant terms (brpa_thc_1b2g) of the THC fragment o=36,
    v=300, NAux=1800, NGridTHC=4032: R from the THC factors, amplitudes
    (NGuess = 3 NAux, EIG cutoff 1e-12) on the GPU, then the timed 1b2g call
    with Zgh = Zgk Zgk^T, the factors and the amplitudes in HBM.  CPU baseline:
    the numpy oracle on 16 of the amplitude vectors, scaled to NT2."""
    import math

    import numpy as np
    import torch

    import brpa
    from brpa import synth
    from oracle import thc as othc
    _one_gpu(a)
    o, v, n, g, nf = synth.THC_CONFIG
    f = synth.make_thc_fragment(o, v, n, g, synth.BASE_SEED)
    x, w, _ = brpa.freq_grid(f.eocc, f.evir, nf)
    d = lambda t: torch.from_numpy(np.ascontiguousarray(t)).cuda()  # noqa: E731
    ctx = brpa.Context(0)
    _, R = ctx.thc_piu(d(f.Xga), d(f.Xgi), d(f.Zgk), f.eocc, f.evir, x, want_R=True)
    A, Vt, _, _, _ = ctx.thc_t2(R, f.eocc, f.evir, x, w, nguess=math.ceil(3.0 * n), seed=1,
                                t2_cutoff_thresh=1e-12)
    nt = len(A)
    Zgh = d(f.Zgk.T @ f.Zgk)            # Zgk (G, N) as math: Zgh = Zgk Zgk^T
    Xi, Ya, U, Ad = d(f.Xgi), d(f.Xga), d(Vt), d(A)
    for _ in range(a.warmup):
        ctx.thc_1b2g(Zgh, Xi, Ya, U, Ad)
    torch.cuda.synchronize()
    t = time.perf_counter()
    kms = 0.0
    for _ in range(a.steps):
        e1, e2, _, _, _ = ctx.thc_1b2g(Zgh, Xi, Ya, U, Ad)
        kms += ctx.last_kernel_ms()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / a.steps
    kms /= a.steps
    flops = 2.0 * g * g * o * nt + 2.0 * g * v * o * nt      # M tiles (both orders) + YU
    tf = flops / (kms * 1e-3) / 1e12
    cpu = None
    if a.cpu_seconds > 0:
        m = 16
        Uh = Vt[:m].reshape(m, o, v).transpose(2, 1, 0)
        Zh = f.Zgk.T @ f.Zgk
        t = time.perf_counter()
        othc.thc_1b2g(Zh, f.Xgi.T, f.Xga.T, Uh, A[:m])
        cpu = (time.perf_counter() - t) * nt / m
    print("# THC 1b2g: NT2 %d, %.1f ms/step, device %.1f ms, %.1f TF/s, Ec1b/2 %.8f Ec2g %.8f"
          % (nt, dt * 1e3, kms, tf, 0.5 * e1, e2), file=sys.stderr, flush=True)
    out = {"metric": "THC SOSEX + 2g cumulant terms of one fragment o=36, v=300, NAux=1800, "
                     "NGridTHC=4032, NT2=%d" % nt,
           "value": round(dt * 1e3, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded PCG64, brpa.synth.make_thc_fragment)",
           "config": {"workload": "THC 1b2g o=36, v=300, NAux=1800, NGridTHC=4032, NT2=%d" % nt},
           "Ec1b_half": 0.5 * e1, "Ec2g": e2,
           "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4),
                        "traffic": None, "kernel": "k_thc_1b2g (+ YU dgemm)",
                        "kernel_ms_per_step": round(kms, 3)},
           "cpu_baseline": None if cpu is None else {
               "value": round(cpu * 1e3, 1), "unit": "ms", "cores": _cpu_threads(),
               "kind": "port", "sample": "numpy oracle on 16 of %d amplitude vectors, scaled" % nt}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="split_i8", choices=list(PRECISIONS))
    ap.add_argument("--nfrag", type=int, default=256)
    ap.add_argument("--nfreqs", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU oracle sample budget (0 disables the CPU baseline)")
    ap.add_argument("--compare-fp64", type=int, default=1,
                    help="also time the other forms (reported under `forms`)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="run brpa_autotune on the batch before timing (split_i8)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="replay the batched launch sequence as a hipGraph (1) or launch eagerly")
    ap.add_argument("--exchange-frags", type=int, default=-1,
                    help="fragments whose T2 + SOSEX is part of each step (default -1: all "
                         "of them; 0 times E_RPA only)")
    ap.add_argument("--exchange-first", type=int, default=0,
                    help="index of the first exchange fragment (the exchange runs on fragments "
                         "first .. first + exchange-frags - 1; a counter-profiling subset)")
    ap.add_argument("--exchange-workers", type=int, default=4,
                    help="worker threads of brpa_exchange_batch (fragments in flight at once)")
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "single", "adjoint", "sosex", "thc", "mbpt3", "aomo",
                             "thc1b2g"],
                    help="batch: the 256-fragment E_RPA + exchange batch (default); single: "
                         "E_RPA of one o=36, v=300, NAux=1800 fragment in every form; "
                         "adjoint: dE/dR and dE/d eps of one o=48, v=400, NAux=2400 "
                         "fragment with a finite-difference check; sosex: the exchange "
                         "contraction of one o=64, v=600, NAux=3200 fragment, FP64 vs "
                         "split-int8; thc: R and Pi(u) of one fragment from THC factors")
    ap.add_argument("--fd-entries", type=int, default=64,
                    help="random R entries in the adjoint workload's finite-difference check")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (reference for the "
                         "FETCH_SIZE/WRITE_SIZE counter scale)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = (1 << 30) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": 8 * n, "copy_bytes_written_per_launch": 8 * n,
                          "launches": 5}))
        return
    if a.workload == "sosex":
        return bench_sosex(a)
    if a.workload == "single":
        return bench_single(a)
    if a.workload == "adjoint":
        return bench_adjoint(a)
    if a.workload == "thc":
        return bench_thc(a)
    if a.workload == "mbpt3":
        return bench_mbpt3(a)
    if a.workload == "aomo":
        return bench_aomo(a)
    if a.workload == "thc1b2g":
        return bench_thc1b2g(a)
    _one_gpu(a)

    import numpy as np
    import torch

    import brpa
    from brpa import synth

    t0 = time.time()
    shapes = synth.batch_shapes(a.nfrag)
    frags = synth.make_batch(a.nfrag, shapes=shapes)
    grids = [brpa.freq_grid(f.eocc, f.evir, a.nfreqs)[:2] for f in frags]
    print("# generated %d fragments in %.1f s" % (a.nfrag, time.time() - t0), file=sys.stderr,
          flush=True)
    dev = torch.device("cuda:0")
    Rd = [torch.from_numpy(f.R).to(dev) for f in frags]
    torch.cuda.synchronize()
    batch = [(Rd[i], f.eocc, f.evir, grids[i][0], grids[i][1]) for i, f in enumerate(frags)]

    ctx = brpa.Context(0)
    tune = None
    if a.autotune and a.precision == "split_i8":
        # the project's autotuner picks the product-kernel variant per NAux bucket
        # on this batch (outside the timed region)
        t = time.perf_counter()
        res = ctx.autotune(batch, "split_i8")
        names = ["T128", "T128_xcd", "T64"]
        tune = {"seconds": round(time.perf_counter() - t, 2),
                "buckets": {str(b): {"choice": names[v], "ms": [round(m, 2) for m in ms]}
                            for b, (v, ms) in res.items()}}
        print("# autotune: %s" % tune, file=sys.stderr, flush=True)

    # the exchange runs in the explicit ov space, FP64, in every form (the split
    # forms cover the Pi products of E_RPA)
    xprec_of = {p: "fp64" for p in PRECISIONS}
    x0 = max(0, min(a.exchange_first, a.nfrag))
    kx = a.nfrag - x0 if a.exchange_frags < 0 else min(a.exchange_frags, a.nfrag - x0)
    xsub = batch[x0:x0 + kx]
    nw = a.exchange_workers

    def erpa_eager(prec):
        # one eager pass: per-stage kernel times (HIP events inside the library)
        ctx.set_graphs(False)
        e = ctx.ecrpa_batch(batch, prec)
        st = ctx.last_stage_ms()
        ctx.set_graphs(bool(a.graphs))
        return e, {k: v[0] for k, v in st.items()}, {k: v[1] for k, v in st.items()}

    def exchange(xprec):
        exs, nts = ctx.exchange_batch(xsub, seed=synth.BASE_SEED + x0, nworkers=nw, precision=xprec)
        stages = {k: v[0] for k, v in ctx.last_exchange_stage_ms().items()}
        return exs, nts, stages

    def step(prec, with_exchange=True):
        """One step: E_RPA of the whole batch (graph replay), then T2 + SOSEX of
        the exchange fragments in this form's exchange arithmetic."""
        torch.cuda.synchronize()
        t0_ = time.perf_counter()
        e = ctx.ecrpa_batch(batch, prec)
        torch.cuda.synchronize()
        t1_ = time.perf_counter()
        x = exchange(xprec_of[prec]) if (with_exchange and kx > 0) else None
        torch.cuda.synchronize()
        t2_ = time.perf_counter()
        return e, x, t1_ - t0_, t2_ - t1_

    # headline form: W warm-up steps, then exactly K timed steps
    p0 = a.precision
    e_eager, stage_ms, stage_n = erpa_eager(p0)
    for _ in range(a.warmup):
        step(p0)
    recs = [step(p0) for _ in range(a.steps)]
    t_erpa = sum(r[2] for r in recs) / a.steps
    t_exch = sum(r[3] for r in recs) / a.steps
    dt = t_erpa + t_exch
    e = recs[-1][0]
    for r in recs:
        if not np.array_equal(r[0], e_eager):
            raise RuntimeError("%s: replayed energies differ from the eager pass" % p0)
    x_last = recs[-1][1]
    if kx > 0:
        for r in recs[:-1]:
            if not (np.array_equal(r[1][0], x_last[0]) and np.array_equal(r[1][1], x_last[1])):
                raise RuntimeError("%s: exchange energies differ between steps" % p0)
    print("# %s: %.1f ms/step (E_RPA %.1f, exchange of %d fragments %.1f) stages %s" %
          (p0, dt * 1e3, t_erpa * 1e3, kx, t_exch * 1e3, stage_ms), file=sys.stderr, flush=True)
    forms = {p0: {"energies": e, "erpa_ms": t_erpa * 1e3, "stage_ms": stage_ms}}
    xres = {}
    if kx > 0:
        xres[xprec_of[p0]] = {"ex": x_last[0], "nt": x_last[1], "stages": x_last[2],
                              "ms": t_exch * 1e3, "steps": a.steps}
    if a.compare_fp64:
        for prec in PRECISIONS:
            if prec in forms:
                continue
            ep, stp, _ = erpa_eager(prec)
            step(prec, with_exchange=False)          # graph capture
            rp = [step(prec, with_exchange=xprec_of[prec] not in xres)
                  for _ in range(max(1, a.steps))]
            if not all(np.array_equal(r[0], ep) for r in rp):
                raise RuntimeError("%s: replayed energies differ from the eager pass" % prec)
            forms[prec] = {"energies": ep, "erpa_ms": sum(r[2] for r in rp) / len(rp) * 1e3,
                           "stage_ms": stp}
            if kx > 0 and xprec_of[prec] not in xres:
                xr = [r for r in rp if r[1] is not None]
                xres[xprec_of[prec]] = {"ex": xr[-1][1][0], "nt": xr[-1][1][1],
                                        "stages": xr[-1][1][2],
                                        "ms": sum(r[3] for r in xr) / len(xr) * 1e3,
                                        "steps": len(xr)}
            print("# %s: E_RPA %.1f ms/step" % (prec, forms[prec]["erpa_ms"]), file=sys.stderr,
                  flush=True)
    for p, d in forms.items():
        d["step_ms"] = d["erpa_ms"] + (xres[xprec_of[p]]["ms"] if kx > 0 else 0.0)

    tot_syrk = sum(syrk_flops(o, v, n, a.nfreqs) for (o, v, n) in shapes)
    tot_chol = sum(chol_flops(n, a.nfreqs) for (o, v, n) in shapes)
    dom = max(("pi", "chol"), key=lambda k: stage_ms[k])
    if dom == "pi":
        kernel, peak = PI_KERNEL[a.precision]
        achieved = tot_syrk / (stage_ms["pi"] * 1e-3) / 1e12
    else:
        kernel, peak = "k_potrf_diag+k_trsm+k_chol_update", PEAK_FP64_TFLOPS
        achieved = tot_chol / (stage_ms["chol"] * 1e-3) / 1e12
    # the committed PMC runs cover the default batch (256 fragments, 16 frequencies)
    traffic = (pmc_traffic(kernel + "<128>")
               if kernel == "k_syrk_i8" and a.nfrag == 256 and a.nfreqs == 16 else None)
    roofline_erpa = {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 3),
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "traffic_source": ("profiles/erpa_batch256_split_i8_kernels.md (FETCH_SIZE x 2.0 "
                                   "+ WRITE_SIZE x 1.0, copy-calibrated)") if traffic else None,
                "kernel": kernel, "kernel_ms_per_step": round(stage_ms[dom], 3),
                "launches_per_step": stage_n[dom]}

    exch = None
    roofline_exch = None
    if kx > 0:
        xp = xprec_of[p0]
        exch = {"n_fragments": kx, "first_fragment": x0, "workers": nw,
                "ms_per_step": round(xres[xp]["ms"], 3),
                "ms_per_fragment": round(xres[xp]["ms"] / kx, 3),
                "ms_per_fragment_by_precision": {q: round(r["ms"] / kx, 3)
                                                 for q, r in xres.items()},
                "stage_ms_per_fragment_by_precision": {
                    q: {k: round(v / kx, 3) for k, v in r["stages"].items()}
                    for q, r in xres.items()},
                "precision": "fp64 (explicit ov-space amplitudes and exchange contraction)",
                "nguess": "ceil(3.0 N), not capped at o*v",
                "n_oversampled": int(sum(1 for (o, v, n) in shapes[x0:x0 + kx]
                                         if o * v <= int(np.ceil(3.0 * n)))),
                "mean_nt2": float(np.mean(xres[xp]["nt"]))}
        if "fp64" in xres and "split_i8" in xres:
            exch["max_abs_dEx_split_i8_vs_fp64_Ha"] = float(
                np.max(np.abs(xres["split_i8"]["ex"] - xres["fp64"]["ex"])))
            exch["n_nt2_differ_split_i8_vs_fp64"] = int(
                np.sum(xres["split_i8"]["nt"] != xres["fp64"]["nt"]))
        # the FP64 MFMA GEMM of the explicit-space exchange (I, I Omega, Gram
        # matrices, triangular-inverse levels, T^T I T, T2), timed in the library
        # (HIP events around every launch, on the worker's stream) over a
        # one-worker pass of fragments 9..16 of the batch, so that its launches
        # do not overlap with another fragment's; the committed counter runs
        # (profiles/exchange_frags9-16_gemm_f64_pmc.md) cover the same fragments
        from brpa import _lib as blib
        r0, kr = (9, 8) if a.nfrag >= 17 else (0, min(16, a.nfrag))
        ctx.exchange_batch(batch[r0:r0 + kr], seed=synth.BASE_SEED + r0, nworkers=1,
                           precision="fp64")
        kst = {name: ctx.last_kernel_stats(kid) for name, kid in
               (("k_gemm_f64", blib.KERNEL_GEMM_F64), ("k_trinv_diag", blib.KERNEL_TRINV_F64),
                ("k_sosex_contract", blib.KERNEL_SOSEX_DOT))}
        gms, gfl, gn = kst["k_gemm_f64"]
        ach = gfl / (gms * 1e-3) / 1e12
        xtraffic = (pmc_traffic("k_gemm_f64 (all variants)",
                                profile="exchange_frags9-16_gemm_f64_pmc.md")
                    if (a.nfrag == 256 and a.nfreqs == 16 and r0 == 9) else None)
        roofline_exch = {
            "bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": xtraffic,
            "traffic_source": ("profiles/exchange_frags9-16_gemm_f64_pmc.md (FETCH_SIZE x 2.0 + "
                               "WRITE_SIZE x 1.0 per dispatch, copy-calibrated, same fragments)")
            if xtraffic else None,
            "kernel": "k_gemm_f64 (FP64 MFMA GEMM/SYRK of the explicit-space exchange)",
            "kernel_ms_per_launch": round(gms / max(gn, 1), 4), "launches": gn,
            "sample": "one-worker exchange pass over fragments %d..%d" % (r0, r0 + kr - 1),
            "other_kernels": {name: {"ms": round(ms, 3), "launches": n,
                                     "tflops": round(fl / max(ms, 1e-9) / 1e9, 3)}
                              for name, (ms, fl, n) in kst.items()}}
        print("# exchange kernels (one worker, %d fragments): %s" % (kr, roofline_exch["other_kernels"]),
              file=sys.stderr, flush=True)

    cpu = None
    parity = None
    if a.cpu_seconds > 0:
        from oracle import erpa as oerpa
        from oracle import sosex as osx
        from oracle import t2 as ot2
        threads = _cpu_threads()
        t = time.perf_counter()
        k = 0
        dmax = {p: 0.0 for p in forms}
        dxmax = {q: 0.0 for q in xres}
        while k < a.nfrag and (k == 0 or time.perf_counter() - t < a.cpu_seconds):
            f = frags[k]
            g = f.gaps()
            x, w = grids[k]
            Pis = [oerpa.gpiug_mo(f.R, g, u) for u in x]
            eref = oerpa.ecrpa_analytic(Pis, x, w)
            for p in forms:
                dmax[p] = max(dmax[p], abs(forms[p]["energies"][k] - eref))
            if x0 <= k < x0 + kx:
                A, V = ot2.t2_v5(f.R, g, x, w, Pis, seed=synth.BASE_SEED + k)
                xref = osx.sosex(f.R, V, A, f.nocc, f.nvir)
                for q in xres:
                    dxmax[q] = max(dxmax[q], abs(xres[q]["ex"][k - x0] - xref))
            k += 1
        tc = time.perf_counter() - t
        cpu = {"value": round(k / tc, 4), "unit": "fragments/sec", "cores": threads,
               "kind": "port", "host": _host_cpus(),
               "sample": "first %d of the %d fragments, E_RPA%s by the numpy FP64 oracle "
                         "(dgemm Pi + dgeqrf log-det%s) on %d threads, %.1f s" %
                         (k, a.nfrag, " + exchange" if kx > 0 else "",
                          "; T2 v5 with the NG x NG Gram eigensolve, SOSEX pair loop"
                          if kx > 0 else "", threads, tc)}
        parity = {"n_checked": k, "max_abs_dE_Ha": dmax[p0], "max_abs_dE_Ha_by_form": dmax,
                  "max_abs_dEx_Ha_by_precision": dxmax if kx > 0 else None}

    value = a.nfrag / (forms[p0]["step_ms"] * 1e-3) if kx == a.nfrag else a.nfrag / t_erpa
    out = {
        "metric": "fragments/sec (E_RPA+exchange corr) split-BF16 vs FP64 at |ΔE|≤1e-7 Ha",
        "value": round(value, 4),
        "unit": "fragments/sec",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(forms[p0]["step_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[p0] + ("; exchange f64" if kx > 0 and p0 != "fp64" else ""),
        "data": "synthetic (seeded PCG64, brpa.synth)",
        "config": {"workload": "batch of %d uneven fragments, o~U{8..40}, v~U{70..350}, "
                               "N=round(4.5(o+v)), %d frequencies" % (a.nfrag, a.nfreqs),
                   "precision": p0,
                   "components": ("E_RPA (batched, %s) + exchange of %d fragments (T2 "
                                  "amplitudes + SOSEX, %s contractions)" %
                                  (p0, kx, xprec_of[p0])) if kx > 0 else "E_RPA only",
                   "step": ("E_RPA of all %d fragments (hipGraph replay) then T2 + SOSEX of "
                            "%s, inputs resident in HBM" %
                            (a.nfrag, "all of them" if kx == a.nfrag else "the first %d" % kx))
                           if kx > 0 else "E_RPA of all fragments",
                   "graphs": bool(a.graphs), "autotune": tune,
                   "algorithmic_tflop_per_erpa_pass": round((tot_syrk + tot_chol) / 1e12, 3)},
        "erpa_only_fragments_per_sec": round(a.nfrag / t_erpa, 4),
        "exchange": exch,
        "forms": {p: {"erpa_fragments_per_sec": round(a.nfrag / (d["erpa_ms"] * 1e-3), 4),
                      "with_exchange_fragments_per_sec": (round(a.nfrag / (d["step_ms"] * 1e-3),
                                                                4) if kx == a.nfrag else None),
                      "erpa_ms_per_step": round(d["erpa_ms"], 3),
                      "step_ms": round(d["step_ms"], 3),
                      "exchange_precision": xprec_of[p] if kx > 0 else None,
                      "stage_ms": {k: round(v, 3) for k, v in d["stage_ms"].items()}}
                  for p, d in forms.items()},
        "roofline": roofline_exch if roofline_exch else roofline_erpa,
        "roofline_erpa": roofline_erpa,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
