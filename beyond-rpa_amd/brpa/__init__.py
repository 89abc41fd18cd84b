"""brpa -- MI355X (gfx950) correlation-energy hot path of beyond-RPA.

Thin Python layer over the C ABI of libbrpa.so (include/brpa.h).  Arrays may
be numpy (host) or torch CUDA tensors (device, float64, contiguous); the
library detects device pointers and otherwise stages through its own device
workspace.  There is no CPU fallback: without the built library every call
raises.

Layouts (bit-identical to the reference's Fortran arrays):
  R[ai][P]  (o*v, N)  == Fortran CRG(P, ai) / Rkai(P, a, i),  ai = i*v + a
  Pi[u][Q][P]         == Fortran PiU(P, Q, u)
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP64, SPLIT_BF16, SPLIT_I8  # noqa: F401

PRECISIONS = {"fp64": FP64, "split_bf16": SPLIT_BF16, "split_i8": SPLIT_I8}


class BrpaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("brpa error %d: %s" % (code, msg))
        self.code = code


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        if x.dtype.itemsize != 8 or not x.is_contiguous():
            raise ValueError("torch tensors must be contiguous float64")
        return C.c_void_p(x.data_ptr())
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("numpy arrays must be C-contiguous float64")
    return C.c_void_p(x.ctypes.data)


def _f64(x):
    if _is_torch(x):
        return x.contiguous()
    return np.ascontiguousarray(x, dtype=np.float64)


def _prec(p):
    if isinstance(p, str):
        return PRECISIONS[p]
    return int(p)


def freq_grid(eocc, evir, nfreqs, core_orb_thresh=_lib.CORE_ORB_THRESH, dai_max_thresh=None):
    """Fixed-size Casimir-Polder grid of one fragment (host code in the library).

    eocc: all occupied energies ascending (those below core_orb_thresh are
    frozen); dai_max_thresh: cap of the gap histogram (None: none).  Returns
    (freqs, weights, info) with info = dict(alpha, rmsd, max_rel_error).
    """
    lib = _lib.load()
    eocc = np.ascontiguousarray(eocc, np.float64)
    evir = np.ascontiguousarray(evir, np.float64)
    x = np.zeros(nfreqs)
    w = np.zeros(nfreqs)
    info = np.zeros(3)
    rc = lib.brpa_freq_grid_ex(_ptr(eocc), eocc.size, _ptr(evir), evir.size, nfreqs,
                               float(core_orb_thresh),
                               0.0 if dai_max_thresh is None else float(dai_max_thresh),
                               _ptr(x), _ptr(w), _ptr(info))
    if rc != 0:
        raise BrpaError(rc, "brpa_freq_grid_ex failed")
    return x, w, dict(alpha=info[0], rmsd=info[1], max_rel_error=info[2])


def optimize_freq_grid(eocc, evir, target_error=1e-6, target_rel_error=1e-3, min_nfreqs=8,
                       max_nfreqs=64):
    lib = _lib.load()
    eocc = np.ascontiguousarray(eocc, np.float64)
    evir = np.ascontiguousarray(evir, np.float64)
    x = np.zeros(max_nfreqs)
    w = np.zeros(max_nfreqs)
    info = np.zeros(3)
    n = C.c_int64(0)
    rc = lib.brpa_optimize_freq_grid(_ptr(eocc), eocc.size, _ptr(evir), evir.size, target_error,
                                     target_rel_error, min_nfreqs, max_nfreqs, _ptr(x), _ptr(w),
                                     C.byref(n), _ptr(info))
    if rc != 0:
        raise BrpaError(rc, "brpa_optimize_freq_grid failed")
    k = n.value
    return x[:k].copy(), w[:k].copy(), dict(alpha=info[0], rmsd=info[1], max_rel_error=info[2],
                                            nfreqs=k)


def dai_histogram(eocc, evir, core_orb_thresh=_lib.CORE_ORB_THRESH, dai_max_thresh=None):
    """Gap histogram of one spin of one system (rpa_DaiHistogram): (values, weights)."""
    lib = _lib.load()
    eocc = np.ascontiguousarray(eocc, np.float64)
    evir = np.ascontiguousarray(evir, np.float64)
    v = np.zeros(_lib.HISTOGRAM_NBINS)
    w = np.zeros(_lib.HISTOGRAM_NBINS)
    rc = lib.brpa_dai_histogram(_ptr(eocc), eocc.size, _ptr(evir), evir.size,
                                float(core_orb_thresh),
                                0.0 if dai_max_thresh is None else float(dai_max_thresh),
                                _ptr(v), _ptr(w))
    if rc != 0:
        raise BrpaError(rc, "brpa_dai_histogram failed")
    return v, w


def dai_max_thresh(spins, core_orb_thresh=_lib.CORE_ORB_THRESH):
    """Largest active gap of a whole complex over its spins [(eocc, evir), ...]
    (rpa_DaiMaxThresh), the cap for its subsystems' histograms."""
    lib = _lib.load()
    eo = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64) for a, _ in spins]))
    ev = np.ascontiguousarray(np.concatenate([np.asarray(b, np.float64) for _, b in spins]))
    no = np.array([len(a) for a, _ in spins], dtype=np.int64)
    nv = np.array([len(b) for _, b in spins], dtype=np.int64)
    t = C.c_double(0.0)
    rc = lib.brpa_dai_max_thresh(len(spins), _ptr(eo), C.c_void_p(no.ctypes.data), _ptr(ev),
                                 C.c_void_p(nv.ctypes.data), float(core_orb_thresh), C.byref(t))
    if rc != 0:
        raise BrpaError(rc, "brpa_dai_max_thresh failed")
    return t.value


def _hist_sets(values, dweights):
    v = np.ascontiguousarray(np.atleast_2d(values), np.float64)
    w = np.ascontiguousarray(np.atleast_2d(dweights), np.float64)
    if v.shape != w.shape or v.shape[1] != _lib.HISTOGRAM_NBINS:
        raise ValueError("histograms must be (nsets, %d) arrays" % _lib.HISTOGRAM_NBINS)
    return v, w


def freq_grid_sets(values, dweights, nfreqs):
    """Grid of nfreqs points on one or several histograms (one per row), the
    worst set deciding the error (rpa_OptQuadScaling)."""
    lib = _lib.load()
    v, w = _hist_sets(values, dweights)
    x = np.zeros(nfreqs)
    fw = np.zeros(nfreqs)
    info = np.zeros(3)
    rc = lib.brpa_opt_quad_scaling(nfreqs, v.shape[0], _ptr(v), _ptr(w), _ptr(x), _ptr(fw),
                                   _ptr(info))
    if rc != 0:
        raise BrpaError(rc, "brpa_opt_quad_scaling failed")
    return x, fw, dict(alpha=info[0], rmsd=info[1], max_rel_error=info[2])


def optimize_freq_quad(values, dweights, target_error=1e-6, target_rel_error=1e-3, min_nfreqs=8,
                       max_nfreqs=64):
    """Adaptive grid on one or several histograms (rpa_OptimizeFreqQuad)."""
    lib = _lib.load()
    v, w = _hist_sets(values, dweights)
    x = np.zeros(max_nfreqs)
    fw = np.zeros(max_nfreqs)
    info = np.zeros(3)
    n = C.c_int64(0)
    rc = lib.brpa_optimize_freq_quad(v.shape[0], _ptr(v), _ptr(w), target_error, target_rel_error,
                                     min_nfreqs, max_nfreqs, _ptr(x), _ptr(fw), C.byref(n),
                                     _ptr(info))
    if rc != 0:
        raise BrpaError(rc, "brpa_optimize_freq_quad failed")
    k = n.value
    return x[:k].copy(), fw[:k].copy(), dict(alpha=info[0], rmsd=info[1],
                                             max_rel_error=info[2], nfreqs=k)


class Context:
    """One device, one HIP stream; not shared between threads."""

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = C.c_void_p()
        rc = self._lib.brpa_ctx_create(device, C.byref(h))
        if rc != 0:
            raise BrpaError(rc, "brpa_ctx_create(device=%d) failed" % device)
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._lib.brpa_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- plumbing --------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self._lib.brpa_last_error(self._h)
            raise BrpaError(rc, msg.decode() if msg else "")

    def _sync_stream(self, *arrays):
        """Run on torch's current stream when torch tensors are involved."""
        if any(_is_torch(a) for a in arrays if a is not None):
            import torch
            s = torch.cuda.current_stream().cuda_stream
            self._lib.brpa_ctx_set_stream(self._h, C.c_void_p(s))
        else:
            self._lib.brpa_ctx_set_stream(self._h, None)

    def last_kernel_ms(self):
        return self._lib.brpa_last_kernel_ms(self._h)

    def last_stage_ms(self):
        """{stage: (ms, launches)} of the last E_RPA call."""
        names = {"pi": _lib.STAGE_PI, "split": _lib.STAGE_SPLIT, "chol": _lib.STAGE_CHOL,
                 "aux": _lib.STAGE_AUX}
        return {k: (self._lib.brpa_last_stage_ms(self._h, v),
                    self._lib.brpa_last_stage_launches(self._h, v)) for k, v in names.items()}

    def last_exchange_stage_ms(self):
        """{stage: (ms, calls)} of the last amplitude, SOSEX or exchange-batch
        call (device time per stage, summed over the fragments of a batch)."""
        names = {"t2_pi": _lib.STAGE_T2_PI, "t2_iomega": _lib.STAGE_T2_IOMEGA,
                 "t2_basis": _lib.STAGE_T2_BASIS, "t2_tit": _lib.STAGE_T2_TIT,
                 "t2_core": _lib.STAGE_T2_CORE, "sosex": _lib.STAGE_SOSEX,
                 "t2_basis_eig": _lib.STAGE_T2_EIG, "t2_basis_drop": _lib.STAGE_T2_DROP}
        return {k: (self._lib.brpa_last_stage_ms(self._h, v),
                    self._lib.brpa_last_stage_launches(self._h, v)) for k, v in names.items()}

    def last_gemm_i8_stats(self):
        """(ms, FP64-equivalent flop, launches) of the split-int8 GEMM launches of
        the last amplitude or exchange-batch call (HIP events around each launch)."""
        ms, fl, n = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
        self._check(self._lib.brpa_last_gemm_i8_stats(self._h, C.byref(ms), C.byref(fl),
                                                      C.byref(n)))
        return ms.value, fl.value, n.value

    def last_kernel_stats(self, kernel):
        """(ms, algorithmic flop, launches) of one measured kernel (brpa._lib.KERNEL_*)
        over the last measured call (HIP events around each launch)."""
        ms, fl, n = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
        self._check(self._lib.brpa_last_kernel_stats(self._h, int(kernel), C.byref(ms),
                                                     C.byref(fl), C.byref(n)))
        return ms.value, fl.value, n.value

    def trim(self):
        self._check(self._lib.brpa_ctx_trim(self._h))

    def set_graphs(self, enable=True):
        """Capture/replay the batched E_RPA launch sequence as a hipGraph."""
        self._check(self._lib.brpa_ctx_set_graphs(self._h, int(bool(enable))))

    def graph_stats(self):
        """(captures, replays) of the last ecrpa_batch call."""
        c, r = C.c_int(0), C.c_int(0)
        self._check(self._lib.brpa_last_graph_stats(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    # -- reference-equivalent entry points -------------------------------
    def gpiug_mo(self, gpiug, R, eocc, evir, freq, prefactor=4.0):
        """gpiug += prefactor * R^T diag(d/(d^2+u^2)) R in place (rpa_GPiUG_MO)."""
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        if nai != eocc.size * evir.size or tuple(gpiug.shape) != (naux, naux):
            raise ValueError("shape mismatch")
        self._sync_stream(gpiug, R)
        self._check(self._lib.brpa_gpiug_mo(self._h, _ptr(gpiug), _ptr(R), naux, _ptr(eocc),
                                            eocc.size, _ptr(evir), evir.size, float(freq),
                                            float(prefactor)))
        return gpiug

    def ecrpa_analytic(self, piu, freqs, weights):
        """E_RPA from stored Pi(u) (rpa_CC_EcRPA_Analytic).  Returns (E, E(u))."""
        nf, n, n2 = piu.shape
        if n != n2:
            raise ValueError("Pi(u) must be square")
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        e = np.zeros(1)
        eu = np.zeros(nf)
        self._sync_stream(piu)
        self._check(self._lib.brpa_ecrpa_analytic(self._h, _ptr(e), _ptr(piu), _ptr(freqs),
                                                  _ptr(weights), nf, n, _ptr(eu)))
        return float(e[0]), eu

    def ecrpa_analytic_noalloc(self, piu, freqs, weights):
        """brpa_ecrpa_analytic_noalloc: the same energy with no device allocation
        (buffers in the library's fixed per-GPU scratch array).  Returns (E, E(u))."""
        nf, n, n2 = piu.shape
        if n != n2:
            raise ValueError("Pi(u) must be square")
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        e = np.zeros(1)
        eu = np.zeros(nf)
        self._sync_stream(piu)
        self._check(self._lib.brpa_ecrpa_analytic_noalloc(self._h, _ptr(e), _ptr(piu),
                                                          _ptr(freqs), _ptr(weights), nf, n,
                                                          _ptr(eu)))
        return float(e[0]), eu

    _BASES = {"random": 1, "eigen": 2, "full": 3, "full_cholesky": 3}

    def rpa_g(self, RWR, basis="eigen", target_error=10 ** -2.5, seed=0):
        """Basis of the effective dielectric matrix (rpa_G).  RWR: (N, N) Pi(u) at
        the lowest frequency, Fortran order (symmetric, so either).  Returns
        (G, RWR', NVecs, AbsError) with G, RWR' as Fortran-ordered (N, N) arrays,
        i.e. C arrays whose [j][i] is element (i, j)."""
        b = self._BASES[basis] if isinstance(basis, str) else int(basis)
        n = RWR.shape[0]
        R = RWR.clone() if _is_torch(RWR) else np.array(RWR, dtype=np.float64, copy=True)
        if _is_torch(R):
            import torch
            G = torch.zeros((n, n), dtype=torch.float64, device=R.device)
        else:
            G = np.zeros((n, n))
        nv, err = C.c_int64(0), C.c_double(0.0)
        self._sync_stream(R, G)
        self._check(self._lib.brpa_rpa_g(self._h, b, _ptr(G), _ptr(R), n, float(target_error),
                                         int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(nv),
                                         C.byref(err)))
        return G, R, nv.value, err.value

    def pack_pi(self, piu, packed=None, accumulate=False):
        """Lower-packed storage (F, N(N+1)/2) of Pi(u) (F, N, N) (rpa_Pack)."""
        nf, n, _ = piu.shape
        if packed is None:
            packed = np.zeros((nf, n * (n + 1) // 2))
        self._sync_stream(piu, packed)
        self._check(self._lib.brpa_pack_pi(self._h, _ptr(packed), _ptr(piu), nf, n,
                                           int(bool(accumulate))))
        return packed

    def unpack_pi(self, packed, n, out=None):
        """Full symmetric Pi(u) (F, N, N) from lower-packed storage (rpa_Unpack)."""
        nf = packed.shape[0]
        if out is None:
            out = np.zeros((nf, n, n))
        self._sync_stream(packed, out)
        self._check(self._lib.brpa_unpack_pi(self._h, _ptr(out), _ptr(packed), nf, n))
        return out

    def ecrpa_packed(self, packed, n, freqs, weights):
        """E_RPA from lower-packed Pi(u) (the energy loop of rpa_FreqIntegral_MO).
        Returns (E, E(u))."""
        nf = packed.shape[0]
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        e = np.zeros(1)
        eu = np.zeros(nf)
        self._sync_stream(packed)
        self._check(self._lib.brpa_ecrpa_packed(self._h, _ptr(e), _ptr(packed), _ptr(freqs),
                                                _ptr(weights), nf, n, _ptr(eu)))
        return float(e[0]), eu

    def ecrpa(self, R, eocc, evir, freqs, weights, precision="fp64"):
        """Fused E_RPA of one fragment.  Returns (E, E(u))."""
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        if nai != eocc.size * evir.size or freqs.size != weights.size:
            raise ValueError("shape mismatch")
        e = np.zeros(1)
        eu = np.zeros(freqs.size)
        self._sync_stream(R)
        self._check(self._lib.brpa_ecrpa(self._h, _prec(precision), _ptr(R), naux, _ptr(eocc),
                                         eocc.size, _ptr(evir), evir.size, _ptr(freqs),
                                         _ptr(weights), freqs.size, _ptr(e), _ptr(eu)))
        return float(e[0]), eu

    def _frag_array(self, frags):
        keep = []
        arr = (_lib.brpa_frag_t * len(frags))()
        for i, (R, eocc, evir, freqs, weights) in enumerate(frags):
            eocc = np.ascontiguousarray(eocc, np.float64)
            evir = np.ascontiguousarray(evir, np.float64)
            freqs = np.ascontiguousarray(freqs, np.float64)
            weights = np.ascontiguousarray(weights, np.float64)
            keep.append((R, eocc, evir, freqs, weights))
            nai, naux = R.shape
            if nai != eocc.size * evir.size or freqs.size != weights.size:
                raise ValueError("fragment %d: shape mismatch" % i)
            arr[i] = _lib.brpa_frag_t(eocc.size, evir.size, naux, freqs.size, _ptr(R).value,
                                      _ptr(eocc).value, _ptr(evir).value, _ptr(freqs).value,
                                      _ptr(weights).value)
        return arr, keep

    def autotune(self, frags, precision="split_i8"):
        """Time the product-kernel variants on these fragments; returns
        {bucket: (chosen variant, [ms per variant])} for the buckets present."""
        arr, keep = self._frag_array(frags)
        self._sync_stream(*[k[0] for k in keep])
        self._check(self._lib.brpa_autotune(self._h, _prec(precision), len(frags), arr))
        out = {}
        for b in range(3):
            v = C.c_int(-1)
            ms = np.zeros(3)
            self._check(self._lib.brpa_autotune_result(self._h, b, C.byref(v), _ptr(ms)))
            if v.value >= 0:
                out[b] = (v.value, ms.tolist())
        return out

    def exchange_batch(self, frags, lam=1.0, cutoff=1e-10, guess_factor=3.0, seed=0,
                       alpha=1.0, s2_operator=False, nworkers=1, precision="fp64"):
        """SOSEX exchange energy of many fragments (T2 then SOSEX per fragment,
        on device-resident amplitudes).  Fragment f uses guess seed seed + f.
        Returns (E_x array, NT2 array)."""
        arr, keep = self._frag_array(frags)
        ex = np.zeros(len(frags))
        nt = np.zeros(len(frags), dtype=np.int64)
        self._sync_stream(*[k[0] for k in keep])
        self._check(self._lib.brpa_exchange_batch(
            self._h, _prec(precision), len(frags), arr, float(lam), float(cutoff),
            float(guess_factor), int(seed) & 0xFFFFFFFFFFFFFFFF, float(alpha),
            int(bool(s2_operator)), int(nworkers), _ptr(ex), C.c_void_p(nt.ctypes.data)))
        return ex, nt

    def ecrpa_batch(self, frags, precision="fp64"):
        """E_RPA of many fragments in batched launches.

        frags: sequence of (R, eocc, evir, freqs, weights).  Returns an array of E.
        """
        arr, keep = self._frag_array(frags)   # keep: the arrays arr points into
        e = np.zeros(len(frags))
        self._sync_stream(*[k[0] for k in keep])
        self._check(self._lib.brpa_ecrpa_batch(self._h, _prec(precision), len(frags), arr,
                                               _ptr(e)))
        return e

    def t2(self, R, eocc, evir, freqs, weights, lam=1.0, cutoff=1e-10, nguess=0, seed=0,
           precision="fp64", device=False):
        """Factorised direct-ring amplitudes (rpa_CC_T2).  Returns (A, V) with
        A ascending (NT2,) and V as C-order V[ai][mu] (ov, NT2); numpy arrays,
        or float64 torch tensors on R's device when device=True."""
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        if nai != eocc.size * evir.size or freqs.size != weights.size:
            raise ValueError("shape mismatch")
        nt = C.c_int64(0)
        self._sync_stream(R)
        self._check(self._lib.brpa_t2_compute(self._h, _prec(precision), _ptr(R), naux, _ptr(eocc),
                                              eocc.size, _ptr(evir), evir.size, _ptr(freqs),
                                              _ptr(weights), freqs.size, float(lam), float(cutoff),
                                              int(nguess), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                              C.byref(nt)))
        k = nt.value
        if device:
            import torch
            dev = R.device if _is_torch(R) else torch.device("cuda", self.device)
            A = torch.zeros(k, dtype=torch.float64, device=dev)
            V = torch.zeros((nai, k), dtype=torch.float64, device=dev)
        else:
            A = np.zeros(k)
            V = np.zeros((nai, k))
        if k:
            self._check(self._lib.brpa_t2_copyout(self._h, _ptr(A), _ptr(V)))
        return A, V

    def thc_piu(self, Xga, Xgi, Zgk, eocc, evir, freqs, prefactor=4.0, PiU=None,
                want_R=False, device=False, out=None, R_out=None):
        """Pi(u) of every frequency from the THC factors (rpa_THC_PiU_BlockAlgo,
        rpa_THC.f90:285-339).  Xga (ngrid, nvir), Xgi (ngrid, nocc), Zgk (ngrid,
        naux) are Fortran-ordered arrays, i.e. C arrays of shape (nvir, ngrid),
        (nocc, ngrid), (naux, ngrid) holding X(g, a) at [a][g].  PiU: an
        existing (nfreqs, naux, naux) array to add to (second spin block), else
        a new one is returned (or `out`, overwritten).  want_R: also return
        R[ai][P] (ov, naux), into `R_out` when given."""
        nvir, ngrid = Xga.shape
        nocc, naux = Xgi.shape[0], Zgk.shape[0]
        if Xgi.shape[1] != ngrid or Zgk.shape[1] != ngrid:
            raise ValueError("shape mismatch")
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        if eocc.size != nocc or evir.size != nvir:
            raise ValueError("shape mismatch")
        nf = freqs.size
        accumulate = PiU is not None
        if accumulate and out is not None:
            raise ValueError("give PiU (add to it) or out (overwrite it), not both")
        torch_out = device or _is_torch(Xga)
        if torch_out:
            import torch
            dev = Xga.device if _is_torch(Xga) else torch.device("cuda", self.device)
            new = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=dev)  # noqa: E731
        else:
            new = lambda *shape: np.zeros(shape)  # noqa: E731
        if PiU is None:
            PiU = out if out is not None else new(nf, naux, naux)
        if tuple(PiU.shape) != (nf, naux, naux):
            raise ValueError("PiU shape mismatch")
        R = None
        if want_R:
            R = R_out if R_out is not None else new(nocc * nvir, naux)
            if tuple(R.shape) != (nocc * nvir, naux):
                raise ValueError("R shape mismatch")
        self._sync_stream(Xga, Xgi, Zgk, PiU)
        self._check(self._lib.brpa_thc_piu(self._h, _ptr(PiU), _ptr(R), _ptr(Xga), _ptr(Xgi),
                                           _ptr(Zgk), ngrid, nocc, nvir, naux, _ptr(eocc),
                                           _ptr(evir), _ptr(freqs), nf, float(prefactor),
                                           int(accumulate)))
        return (PiU, R) if want_R else PiU

    def last_thc_stage_ms(self):
        """{stage: (ms, launches)} of the last thc_piu call."""
        names = {"thc_r": _lib.STAGE_THC_R, "thc_pi": _lib.STAGE_THC_PI}
        return {k: (self._lib.brpa_last_stage_ms(self._h, v),
                    self._lib.brpa_last_stage_launches(self._h, v)) for k, v in names.items()}

    def thc_1b2g(self, Zgh, Xgi, Yga, Uaim, Am, want_diag=False):
        """THC SOSEX + 2g cumulant terms (rpa_CCD_corrections_1b2g).  Fortran-
        ordered inputs as C arrays: Zgh (G, G) symmetric, Xgi (o, G), Yga (v, G),
        Uaim (NT2, o, v) = the (NT2, ov) V of thc_t2, Am (NT2,).  Returns
        (Ec1b, Ec2g, EigSOSEX, Eig2g, YXUggm or None); Ec1b before the 1/2."""
        G = Zgh.shape[0]
        o, v = Xgi.shape[0], Yga.shape[0]
        Am = np.ascontiguousarray(Am, np.float64) if not _is_torch(Am) else Am
        nt = Am.numel() if _is_torch(Am) else Am.size
        nu = Uaim.numel() if _is_torch(Uaim) else Uaim.size
        if nu != nt * o * v or Xgi.shape[1] != G or Yga.shape[1] != G:
            raise ValueError("shape mismatch")
        e1, e2 = np.zeros(1), np.zeros(1)
        es, eg = np.zeros(nt), np.zeros(nt)
        diag = np.zeros((nt, G)) if want_diag else None
        self._sync_stream(Zgh, Xgi, Yga, Uaim, Am)
        self._check(self._lib.brpa_thc_1b2g(self._h, _ptr(Zgh), _ptr(Xgi), _ptr(Yga), _ptr(Uaim),
                                            _ptr(Am), G, o, v, nt, _ptr(e1), _ptr(e2), _ptr(es),
                                            _ptr(eg), _ptr(diag)))
        return float(e1[0]), float(e2[0]), es, eg, diag

    _CUTOFF_TYPES = {"eig": 0, "eig/maxeig": 1, "eig/nelectron": 2, "eig/nocc": 2,
                     "sum-rejected": 3, "sum_rejected": 3}

    def thc_t2(self, R, eocc, evir, freqs, weights, lam=1.0, cutoff=1e-10, nguess=0, seed=0,
               t2_cutoff_thresh=1e-4, t2_cutoff_type="eig", smooth_step=False, steepness=0.1,
               common_thresh=-1.0, precision="fp64"):
        """Amplitudes of the THC route (rpa_THC_CC_T2, rpa_CC_Doubles.f90:645-984):
        brpa_t2_compute, the small-eigenvalue cutoff, then EigRPA.

        Returns (A, V, EigRPA, common_out, n_scaled) with V in the route's
        layout V(ov, NT2), i.e. a C-order (NT2, ov) array V[mu][ai].
        common_thresh < 0: supersystem (common_out = the K it used); otherwise
        the supersystem's K is applied and returned unchanged.
        """
        ctype = (self._CUTOFF_TYPES[t2_cutoff_type.lower()] if isinstance(t2_cutoff_type, str)
                 else int(t2_cutoff_type))
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        if nai != eocc.size * evir.size or freqs.size != weights.size:
            raise ValueError("shape mismatch")
        nt = C.c_int64(0)
        self._sync_stream(R)
        self._check(self._lib.brpa_t2_compute(self._h, _prec(precision), _ptr(R), naux, _ptr(eocc),
                                              eocc.size, _ptr(evir), evir.size, _ptr(freqs),
                                              _ptr(weights), freqs.size, float(lam), float(cutoff),
                                              int(nguess), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                              C.byref(nt)))
        if nt.value == 0:
            raise BrpaError(_lib.ERR_ARG, "no T2 eigenvectors above the Gram cutoff")
        common = C.c_double(float(common_thresh))
        nsc = C.c_int64(0)
        self._check(self._lib.brpa_t2_cutoff(self._h, ctype, float(t2_cutoff_thresh),
                                             int(bool(smooth_step)), float(steepness), eocc.size,
                                             C.byref(common), C.byref(nt), C.byref(nsc)))
        k = nt.value
        A, V, eig = np.zeros(k), np.zeros((k, nai)), np.zeros(k)
        self._check(self._lib.brpa_t2_copyout_ov(self._h, _ptr(A), _ptr(V)))
        self._check(self._lib.brpa_t2_eig_rpa(self._h, _ptr(R), naux, _ptr(eig)))
        return A, V, eig, common.value, nsc.value

    def sosex(self, V, A, R, nocc, nvir, alpha=1.0, s2_operator=False, precision="fp64"):
        """Exchange correction -alpha sum (aj|bi) T(ai,bj) (rpa_CC_Exchange_SOSEX)."""
        nai, naux = R.shape
        if not _is_torch(A):
            A = np.ascontiguousarray(A, np.float64)
        nt = A.numel() if _is_torch(A) else A.size
        if nai != nocc * nvir or tuple(V.shape) != (nai, nt):
            raise ValueError("shape mismatch")
        e = np.zeros(1)
        self._sync_stream(V, R)
        self._check(self._lib.brpa_sosex(self._h, _prec(precision), _ptr(e), float(alpha), _ptr(V),
                                         _ptr(A), _ptr(R), nocc, nvir, naux, nt,
                                         int(bool(s2_operator))))
        return float(e[0])

    # ---- dense FP64 building blocks (column-major, device tensors) ----------
    @staticmethod
    def _fmat(x, name):
        """(pointer, rows, cols, leading dimension) of a column-major (Fortran-
        ordered) 2-D float64 CUDA tensor, e.g. torch.empty(n, m).T for an m x n
        matrix."""
        if not _is_torch(x) or x.dim() != 2 or x.dtype.itemsize != 8 or not x.is_cuda:
            raise ValueError("%s: a 2-D float64 CUDA tensor is required" % name)
        m, n = x.shape
        if x.stride(0) != 1 or (n > 1 and x.stride(1) < max(1, m)):
            raise ValueError("%s: column-major storage is required" % name)
        return C.c_void_p(x.data_ptr()), m, n, max(1, x.stride(1) if n > 1 else m)

    def dgemm(self, A, B, Cm, transa=False, transb=False, alpha=1.0, beta=0.0):
        """Cm = alpha op(A) op(B) + beta Cm (brpa_dgemm), column-major tensors."""
        pa, ra, ca, lda = self._fmat(A, "A")
        pb, rb, cb, ldb = self._fmat(B, "B")
        pc, m, n, ldc = self._fmat(Cm, "C")
        k = ra if transa else ca
        if (ca if transa else ra) != m or (rb if transb else cb) != n or (cb if transb else rb) != k:
            raise ValueError("shape mismatch")
        self._sync_stream(A, B, Cm)
        self._check(self._lib.brpa_dgemm(self._h, int(bool(transa)), int(bool(transb)), m, n, k,
                                         float(alpha), pa, lda, pb, ldb, float(beta), pc, ldc))
        return Cm

    def dsyrk(self, A, Cm, trans=False, alpha=1.0, beta=0.0, fill_lower=True):
        """Upper (and with fill_lower the lower) triangle of Cm = alpha op(A) op(A)^T
        + beta Cm (brpa_dsyrk), column-major tensors."""
        pa, ra, ca, lda = self._fmat(A, "A")
        pc, n, n2, ldc = self._fmat(Cm, "C")
        k = ra if trans else ca
        if n != n2 or (ca if trans else ra) != n:
            raise ValueError("shape mismatch")
        self._sync_stream(A, Cm)
        self._check(self._lib.brpa_dsyrk(self._h, int(bool(trans)), int(bool(fill_lower)), n, k,
                                         float(alpha), pa, lda, float(beta), pc, ldc))
        return Cm

    def dtrsm(self, U, B, side_right=False):
        """B <- U^{-T} B (side_right False) or B U^{-1} (True), U upper triangular
        (brpa_dtrsm), column-major tensors."""
        pu, ru, cu, ldu = self._fmat(U, "U")
        pb, m, n, ldb = self._fmat(B, "B")
        if ru != cu or ru != (n if side_right else m):
            raise ValueError("shape mismatch")
        self._sync_stream(U, B)
        self._check(self._lib.brpa_dtrsm(self._h, int(bool(side_right)), m, n, pu, ldu, pb, ldb))
        return B

    def dtrtri(self, U, W):
        """W_z = U_z^{-1} for upper triangular U (brpa_dtrtri): U and W column-major
        (n, n) CUDA tensors, or batches (nmat, n, n) whose [z] are column-major
        (i.e. C-contiguous (nmat, n, n) holding U_z^T in row-major order)."""
        if U.dim() == 2:
            pu, n, n2, ldu = self._fmat(U, "U")
            pw, nw, nw2, ldw = self._fmat(W, "W")
            nmat, su, sw = 1, 0, 0
        else:
            if not (U.is_cuda and W.is_cuda and U.is_contiguous() and W.is_contiguous()):
                raise ValueError("batched U, W must be contiguous CUDA tensors")
            nmat, n, n2 = U.shape
            nw, nw2 = W.shape[1:]
            pu, pw = _ptr(U), _ptr(W)
            ldu = ldw = n
            su = sw = n * n
        if n != n2 or nw != n or nw2 != n:
            raise ValueError("shape mismatch")
        self._sync_stream(U, W)
        self._check(self._lib.brpa_dtrtri(self._h, n, nmat, pu, ldu, su, pw, ldw, sw))
        return W

    def ao_mo(self, RG, pairs, Co, Cv, out=None, accumulate=False):
        """AO -> MO transform of the Cholesky vectors (rpa_CRG / rpa_C).

        RG (npq, nvecs) C array = Fortran RG(nvecs, npq); pairs (npq, 2) int32
        AO indices of each packed column; Co (nocc, nao) = Fortran (nao, nocc),
        Cv (nvir, nao).  Returns R[ai][P] (ov, nvecs): `out` when given (added
        to when accumulate), else a new array on the inputs' side."""
        npq, nvecs = RG.shape
        nocc, nao = Co.shape
        nvir = Cv.shape[0]
        pairs = np.ascontiguousarray(pairs, np.int32)
        if pairs.shape != (npq, 2) or Cv.shape[1] != nao:
            raise ValueError("shape mismatch")
        if out is None:
            if accumulate:
                raise ValueError("accumulate needs out")
            if _is_torch(RG):
                import torch
                out = torch.zeros((nocc * nvir, nvecs), dtype=torch.float64, device=RG.device)
            else:
                out = np.zeros((nocc * nvir, nvecs))
        if tuple(out.shape) != (nocc * nvir, nvecs):
            raise ValueError("out shape mismatch")
        self._sync_stream(RG, Co, Cv, out)
        self._check(self._lib.brpa_ao_mo(self._h, _ptr(out), _ptr(RG), nvecs, npq,
                                         C.c_void_p(pairs.ctypes.data), _ptr(Co), _ptr(Cv), nao,
                                         nocc, nvir, int(bool(accumulate))))
        return out

    def mbpt3_1(self, V, A, R, eocc, evir, precision="fp64"):
        """Third-order exchange MBPT3-1 (rpa_CC_Exchange_MBPT3_1): Lyapunov
        equation for X = V^T (dT/dlambda) V, then -sum (aj|bi) (V X V^T)(ai,bj)."""
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        if not _is_torch(A):
            A = np.ascontiguousarray(A, np.float64)
        nt = A.numel() if _is_torch(A) else A.size
        if nai != eocc.size * evir.size or tuple(V.shape) != (nai, nt):
            raise ValueError("shape mismatch")
        e = np.zeros(1)
        self._sync_stream(V, R, A)
        self._check(self._lib.brpa_mbpt3_1(self._h, _prec(precision), _ptr(e), _ptr(V), _ptr(A),
                                           nt, _ptr(R), naux, _ptr(eocc), _ptr(evir), eocc.size,
                                           evir.size))
        return float(e[0])

    def mbpt3_2(self, R, V, A, nocc, nvir):
        """Third-order exchange MBPT3-2 (rpa_CC_Exchange_MBPT3_2):
        2 sum_{c,k,l} (sum_{a,i} R(l, ai) T(ak, ci))^2."""
        nai, naux = R.shape
        if not _is_torch(A):
            A = np.ascontiguousarray(A, np.float64)
        nt = A.numel() if _is_torch(A) else A.size
        if nai != nocc * nvir or tuple(V.shape) != (nai, nt):
            raise ValueError("shape mismatch")
        e = np.zeros(1)
        self._sync_stream(V, R, A)
        self._check(self._lib.brpa_mbpt3_2(self._h, _ptr(e), _ptr(R), _ptr(V), _ptr(A), nocc,
                                           nvir, naux, nt))
        return float(e[0])

    def last_mbpt3_stage_ms(self):
        """{stage: (ms, launches)} of the last mbpt3_1 / mbpt3_2 call."""
        names = {"mbpt3": _lib.STAGE_MBPT3, "sosex": _lib.STAGE_SOSEX}
        return {k: (self._lib.brpa_last_stage_ms(self._h, v),
                    self._lib.brpa_last_stage_launches(self._h, v)) for k, v in names.items()}

    def ecrpa_grad(self, R, eocc, evir, freqs, weights):
        """(dE/dR, dE/d eps_occ, dE/d eps_vir) of the FP64 E_RPA, grid held fixed.

        dE/dR has R's type and device (numpy array or torch tensor)."""
        nai, naux = R.shape
        eocc, evir = np.ascontiguousarray(eocc, np.float64), np.ascontiguousarray(evir, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        weights = np.ascontiguousarray(weights, np.float64)
        if nai != eocc.size * evir.size or freqs.size != weights.size:
            raise ValueError("shape mismatch")
        if _is_torch(R):
            import torch
            dR = torch.empty_like(R)
        else:
            dR = np.empty_like(R)
        deo = np.zeros(eocc.size)
        dev = np.zeros(evir.size)
        self._sync_stream(R)
        self._check(self._lib.brpa_ecrpa_grad(self._h, _ptr(R), naux, _ptr(eocc), eocc.size,
                                              _ptr(evir), evir.size, _ptr(freqs), _ptr(weights),
                                              freqs.size, _ptr(dR), _ptr(deo), _ptr(dev)))
        return dR, deo, dev
