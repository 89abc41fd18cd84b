"""The C ABI library loads and exports every function include/brpa.h declares
(no compute calls: this runs on machines without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "brpa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(brpa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("brpa_ctx_create", "brpa_gpiug_mo", "brpa_ecrpa_analytic", "brpa_ecrpa",
              "brpa_ecrpa_batch", "brpa_freq_grid"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from brpa import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_no_gpu_means_a_loud_error():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import brpa
    with pytest.raises(brpa.BrpaError):
        brpa.Context(0)


def test_fortran_shim_exports_reference_routines():
    import shutil
    lib = os.path.join(ROOT, "beyond-rpa_amd", "brpa", "libbrpa_fortran.so")
    if not os.path.exists(lib):
        pytest.skip("Fortran shim not built (amdflang absent)")
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    import subprocess
    out = subprocess.run([nm, "-D", "--defined-only", lib], capture_output=True, text=True).stdout.lower()
    for name in ("rpa_gpiug_mo", "rpa_cc_ecrpa_analytic", "rpa_cc_t2", "rpa_cc_exchange_sosex"):
        assert name in out, name
