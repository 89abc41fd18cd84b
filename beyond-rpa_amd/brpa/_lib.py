"""ctypes binding of libbrpa.so (declarations mirror include/brpa.h)."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbrpa.so")

(OK, ERR_ARG, ERR_HIP, ERR_NOT_SPD, ERR_GAP, ERR_NONFINITE, ERR_OOM, ERR_GRID, ERR_SOLVER,
 ERR_INTERNAL) = range(10)
FP64, SPLIT_BF16, SPLIT_I8 = 0, 1, 2
STAGE_PI, STAGE_SPLIT, STAGE_CHOL, STAGE_AUX = range(4)
# amplitude-eigenvalue cutoff rules (brpa_t2_cutoff)
T2_CUTOFF_EIG, T2_CUTOFF_EIG_DIV_MAXEIG, T2_CUTOFF_EIG_DIV_NELECTRON, T2_CUTOFF_SUM_REJECTED = range(4)
STAGE_T2_PI, STAGE_T2_IOMEGA, STAGE_T2_BASIS, STAGE_T2_TIT, STAGE_T2_CORE, STAGE_SOSEX = range(4, 10)
STAGE_THC_R, STAGE_THC_PI, STAGE_MBPT3, STAGE_T2_EIG, STAGE_T2_DROP = 10, 11, 12, 13, 14
# measured kernels (brpa_last_kernel_stats)
KERNEL_GEMM_I8, KERNEL_GEMM_F64, KERNEL_SOSEX_DOT, KERNEL_TRSM_F64, KERNEL_TRINV_F64 = range(5)
HISTOGRAM_NBINS = 100          # BRPA_HISTOGRAM_NBINS
CORE_ORB_THRESH = -3.0         # BRPA_CORE_ORB_THRESH

_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class brpa_frag_t(C.Structure):
    _fields_ = [("nocc", _i64), ("nvir", _i64), ("naux", _i64), ("nfreqs", _i64),
                ("R", _vp), ("eocc", _vp), ("evir", _vp), ("freqs", _vp), ("weights", _vp)]


# name -> (restype, argtypes)
SIGNATURES = {
    "brpa_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "brpa_ctx_destroy": (C.c_int, [_vp]),
    "brpa_last_error": (C.c_char_p, [_vp]),
    "brpa_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "brpa_ctx_trim": (C.c_int, [_vp]),
    "brpa_ctx_set_graphs": (C.c_int, [_vp, C.c_int]),
    "brpa_last_graph_stats": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "brpa_last_kernel_ms": (C.c_double, [_vp]),
    "brpa_last_stage_ms": (C.c_double, [_vp, C.c_int]),
    "brpa_last_stage_launches": (C.c_int, [_vp, C.c_int]),
    "brpa_last_gemm_i8_stats": (C.c_int, [_vp, _vp, _vp, _vp]),
    "brpa_last_kernel_stats": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "brpa_freq_grid": (C.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp]),
    "brpa_freq_grid_ex": (C.c_int, [_vp, _i64, _vp, _i64, _i64, C.c_double, C.c_double, _vp, _vp,
                                    _vp]),
    "brpa_dai_histogram": (C.c_int, [_vp, _i64, _vp, _i64, C.c_double, C.c_double, _vp, _vp]),
    "brpa_dai_max_thresh": (C.c_int, [_i64, _vp, _vp, _vp, _vp, C.c_double,
                                      C.POINTER(C.c_double)]),
    "brpa_opt_quad_scaling": (C.c_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "brpa_optimize_freq_quad": (C.c_int, [_i64, _vp, _vp, C.c_double, C.c_double, _i64, _i64,
                                          _vp, _vp, C.POINTER(_i64), _vp]),
    "brpa_optimize_freq_grid": (C.c_int, [_vp, _i64, _vp, _i64, C.c_double, C.c_double, _i64,
                                          _i64, _vp, _vp, C.POINTER(_i64), _vp]),
    "brpa_gpiug_mo": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, C.c_double,
                                C.c_double]),
    "brpa_ecrpa_analytic": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "brpa_ecrpa_analytic_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "brpa_ecrpa": (C.c_int, [_vp, C.c_int, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                             _vp, _vp]),
    "brpa_ecrpa_batch": (C.c_int, [_vp, C.c_int, _i64, C.POINTER(brpa_frag_t), _vp]),
    "brpa_autotune": (C.c_int, [_vp, C.c_int, _i64, C.POINTER(brpa_frag_t)]),
    "brpa_autotune_result": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), _vp]),
    "brpa_t2_compute": (C.c_int, [_vp, C.c_int, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                                  C.c_double, C.c_double, _i64, C.c_uint64, C.POINTER(_i64)]),
    "brpa_t2_copyout": (C.c_int, [_vp, _vp, _vp]),
    "brpa_t2_cutoff": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_double, _i64,
                                 C.POINTER(C.c_double), C.POINTER(_i64), C.POINTER(_i64)]),
    "brpa_t2_eig_rpa": (C.c_int, [_vp, _vp, _i64, _vp]),
    "brpa_t2_copyout_ov": (C.c_int, [_vp, _vp, _vp]),
    "brpa_thc_piu": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp,
                               _vp, _i64, C.c_double, C.c_int]),
    "brpa_mbpt3_1": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                               _i64]),
    "brpa_mbpt3_2": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64]),
    "brpa_ao_mo": (C.c_int, [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _i64,
                             C.c_int]),
    "brpa_pack_pi": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.c_int]),
    "brpa_unpack_pi": (C.c_int, [_vp, _vp, _vp, _i64, _i64]),
    "brpa_ecrpa_packed": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "brpa_rpa_g": (C.c_int, [_vp, C.c_int, _vp, _vp, _i64, C.c_double, C.c_uint64,
                             C.POINTER(_i64), C.POINTER(C.c_double)]),
    "brpa_thc_1b2g": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp,
                                _vp, _vp, _vp]),
    "brpa_exchange_batch": (C.c_int, [_vp, C.c_int, _i64, C.POINTER(brpa_frag_t), C.c_double,
                                      C.c_double, C.c_double, C.c_uint64, C.c_double, C.c_int,
                                      C.c_int, _vp, _vp]),
    "brpa_ecrpa_grad": (C.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp,
                                  _vp]),
    "brpa_sosex": (C.c_int, [_vp, C.c_int, _vp, C.c_double, _vp, _vp, _vp, _i64, _i64, _i64,
                             _i64, C.c_int]),
    "brpa_dgemm": (C.c_int, [_vp, C.c_int, C.c_int, _i64, _i64, _i64, C.c_double, _vp, _i64,
                             _vp, _i64, C.c_double, _vp, _i64]),
    "brpa_dsyrk": (C.c_int, [_vp, C.c_int, C.c_int, _i64, _i64, C.c_double, _vp, _i64,
                             C.c_double, _vp, _i64]),
    "brpa_dtrsm": (C.c_int, [_vp, C.c_int, _i64, _i64, _vp, _i64, _vp, _i64]),
    "brpa_dtrtri": (C.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64]),
}

_lib = None


def load():
    """Load libbrpa.so; fail loudly when the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libbrpa.so not found at %s: build it with "
                           "`python beyond-rpa_amd/build.py` (there is no CPU fallback)" % LIB_PATH)
    # torch ships its own libamdhip64 (same SONAME, requested under another
    # name by torch's libraries).  Loading torch first makes libbrpa bind to
    # that runtime, so a process that uses both has one HIP runtime, not two.
    if os.environ.get("BRPA_NO_TORCH", "0") != "1":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
