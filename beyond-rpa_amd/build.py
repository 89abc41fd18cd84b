"""Build the gfx950 shared library libbrpa.so in-tree with hipcc.

    python beyond-rpa_amd/build.py [--jobs N]

Every .hip/.cpp file under csrc/ is compiled for gfx950 and linked into
beyond-rpa_amd/brpa/libbrpa.so, next to the Python wrapper that loads it.
The Fortran shim (fortran/brpa_mod.f90) is compiled with amdflang when it is
available, into brpa/libbrpa_fortran.so.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "brpa")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libbrpa.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CXXFLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]
LDFLAGS = ["-shared", "-fPIC", "-pthread", "--offload-arch=gfx950", "-L", os.path.join(ROCM, "lib"),
           "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def compile_one(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CXXFLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + CXXFLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines()
            if "packed-fp32-ops" not in l and l.strip()]
    return src, obj, r.returncode, "\n".join(msgs)


def build_fortran():
    flang = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")
    src = os.path.join(HERE, "fortran", "brpa_mod.f90")
    if not (os.path.exists(flang) and os.path.exists(src)):
        return None
    out = os.path.join(OUT_DIR, "libbrpa_fortran.so")
    mod_dir = os.path.join(OBJ_DIR, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    cmd = [flang, "-fdefault-integer-8", "-O2", "-fPIC", "-shared", "-module-dir", mod_dir,
           src, "-o", out, "-L", OUT_DIR, "-lbrpa", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + r.stdout + r.stderr)
    for name in ("shim_check", "grid_check"):
        drv = os.path.join(HERE, "fortran", name + ".f90")
        exe = os.path.join(OUT_DIR, "brpa_" + name)
        cmd = [flang, "-fdefault-integer-8", "-O2", "-I", mod_dir, drv, "-o", exe, "-L", OUT_DIR,
               "-lbrpa_fortran", "-lbrpa", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed on %s:\n%s%s" % (name, r.stdout, r.stderr))
    return out


def build(jobs=8, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, msg in ex.map(compile_one, srcs):
            if msg and (verbose or rc != 0):
                print(msg, file=sys.stderr)
            if rc != 0:
                raise RuntimeError("hipcc failed on %s" % src)
            objs.append(obj)
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC] + objs + LDFLAGS + ["-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    fortran = build_fortran()
    return LIB, fortran


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
