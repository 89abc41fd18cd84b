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


def test_fragment_array_points_at_kept_arrays():
    """The brpa_frag_t array handed to the batch entry points points into arrays
    the binding keeps alive through the call, and it rejects what the library
    would misread (ragged shapes, non-contiguous or non-float64 R)."""
    import numpy as np

    import brpa
    from brpa import synth
    f = synth.make_fragment(3, 10, 20, seed=1)
    x = np.linspace(0.1, 1.0, 4)
    frags = [(f.R, f.eocc.tolist(), f.evir, x, np.ones(4)), (f.R, f.eocc, f.evir, x[:3], np.ones(3))]
    arr, keep = brpa.Context._frag_array(None, frags)
    for i, (R, eo, ev, fx, fw) in enumerate(keep):
        a = arr[i]
        assert (a.nocc, a.nvir, a.naux, a.nfreqs) == (3, 10, 20, fx.size)
        assert a.R == R.ctypes.data and a.eocc == eo.ctypes.data and a.evir == ev.ctypes.data
        assert a.freqs == fx.ctypes.data and a.weights == fw.ctypes.data
    with pytest.raises(ValueError):          # o*v rows do not match the energies
        brpa.Context._frag_array(None, [(f.R, f.eocc[:2], f.evir, x, np.ones(4))])
    with pytest.raises(ValueError):          # frequencies and weights differ in length
        brpa.Context._frag_array(None, [(f.R, f.eocc, f.evir, x, np.ones(3))])
    with pytest.raises(ValueError):          # a transposed view is not C-contiguous
        brpa.Context._frag_array(None, [(np.asfortranarray(f.R), f.eocc, f.evir, x, np.ones(4))])
    with pytest.raises(ValueError):
        brpa.Context._frag_array(None, [(f.R.astype(np.float32), f.eocc, f.evir, x, np.ones(4))])


def _prototypes():
    text = open(os.path.join(ROOT, "include", "brpa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return dict(re.findall(r"\b(brpa_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S))


def test_binding_argument_kinds_match_the_header():
    """Every ctypes signature has the header's parameters in order: pointer,
    int64_t, int, double or uint64_t (an int passed where the C side reads
    int64_t would leave the upper half of the register undefined)."""
    import ctypes as C

    from brpa import _lib

    def kind_h(arg):
        arg = arg.strip()
        if "*" in arg:
            return "ptr"
        t = arg.rsplit(None, 1)[0].replace("const ", "").strip()
        return {"int64_t": "i64", "int": "i32", "double": "f64", "uint64_t": "u64"}[t]

    def kind_c(x):
        if x in (C.c_void_p, C.c_char_p) or (isinstance(x, type) and issubclass(x, C._Pointer)):
            return "ptr"
        return {C.c_int64: "i64", C.c_int: "i32", C.c_double: "f64", C.c_uint64: "u64"}[x]

    protos = _prototypes()
    assert set(protos) == set(declared_functions())
    for name, args in protos.items():
        want = [] if args.strip() in ("", "void") else [kind_h(a) for a in args.split(",")]
        have = [kind_c(x) for x in _lib.SIGNATURES[name][1]]
        assert have == want, (name, want, have)
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "brpa.h")).read(), flags=re.S)
    rets = dict((n, r) for r, n in re.findall(r"(int|double|const char \*)\s*(brpa_[a-z0-9_]+)\s*\(", text))
    assert set(rets) == set(protos)
    for name, r in rets.items():
        assert _lib.SIGNATURES[name][0] is {"int": C.c_int, "double": C.c_double,
                                             "const char *": C.c_char_p}[r], name


def test_fortran_interfaces_match_the_header():
    """The bind(C) interfaces of the Fortran module (brpa_mod.f90) pass each
    parameter as the header declares it: by value for int64_t / int / double
    (uint64_t as a 64-bit integer, Fortran has no unsigned kind), by
    reference or as type(c_ptr) for pointers."""
    src = open(os.path.join(ROOT, "beyond-rpa_amd", "fortran", "brpa_mod.f90")).read()
    src = re.sub(r"&\s*\n\s*", " ", src)            # join continuation lines
    blocks = re.findall(r"function\s+\w+\s*\(([^)]*)\)\s*bind\(C,\s*name=\"(brpa_\w+)\"\)(.*?)end function",
                        src, flags=re.S | re.I)
    assert len(blocks) >= 15
    protos = _prototypes()

    def kind_h(arg):
        arg = arg.strip()
        if "*" in arg:
            return "ptr"
        t = arg.rsplit(None, 1)[0].replace("const ", "").strip()
        return {"int64_t": "i64", "uint64_t": "i64", "int": "i32", "double": "f64"}[t]

    for arglist, cname, body in blocks:
        names = [a.strip().lower() for a in arglist.split(",") if a.strip()]
        decl = {}
        for line in body.splitlines():
            if "::" not in line:
                continue
            spec, vars_ = line.split("::", 1)
            spec = spec.strip().lower().replace(" ", "")
            value = ",value" in spec
            for v in re.split(r",(?![^(]*\))", vars_):
                v = v.strip().lower()
                if not v:
                    continue
                v = v.split("(")[0].strip()
                if not value or spec.startswith("type(c_ptr)"):
                    decl[v] = "ptr"
                elif spec.startswith("integer(c_int64_t)"):
                    decl[v] = "i64"
                elif spec.startswith("integer(c_int)"):
                    decl[v] = "i32"
                elif spec.startswith("real(c_double)"):
                    decl[v] = "f64"
        have = [decl[n] for n in names]
        args = protos[cname]
        want = [] if args.strip() in ("", "void") else [kind_h(a) for a in args.split(",")]
        assert have == want, (cname, want, have)
