"""The library's frequency-grid host code (beyond-rpa_amd/csrc/grid.cpp) under
AddressSanitizer and UndefinedBehaviorSanitizer (float-cast-overflow
included), driven by tests/grid_fuzz.cpp over random and degenerate orbital
energies.  A zero-width gap range once cast a NaN bin quotient to an integer
here; the sanitizers stop the run at any such finding."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(HIP_INC),
                    reason="needs g++ and the HIP headers")
def test_grid_host_code_is_sanitizer_clean(tmp_path):
    flags = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined,float-cast-overflow",
             "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__",
             "-I" + os.path.join(ROOT, "include"), "-I" + HIP_INC]
    obj = str(tmp_path / "grid.o")
    exe = str(tmp_path / "grid_fuzz")
    subprocess.run(["g++", *flags, "-c", os.path.join(ROOT, "beyond-rpa_amd", "csrc", "grid.cpp"),
                    "-o", obj], check=True, timeout=300)
    subprocess.run(["g++", *flags, os.path.join(ROOT, "tests", "grid_fuzz.cpp"), obj, "-o", exe],
                   check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, "40"], capture_output=True, text=True, timeout=600, env=env)
    if "does not come first" in r.stderr:
        pytest.skip("the AddressSanitizer runtime cannot load first in this process environment")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    ok, err = (int(x) for x in r.stdout.split()[1::2])
    assert ok > 0 and err > 0      # both the valid and the rejected inputs were exercised
