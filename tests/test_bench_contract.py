"""Host-side pieces of bench.py that need no GPU: the HBM traffic it reads from
the committed counter summaries, the flop counts it prices kernels with, and
the one-GPU refusal of a multi-GPU request."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("kernel,profile", [
    ("k_syrk_i8<128>", "erpa_batch256_split_i8_kernels.md"),
    ("k_gemm_f64 (all variants)", "exchange_frags9-16_gemm_f64_pmc.md"),
])
def test_traffic_rows_of_the_committed_counter_runs(kernel, profile):
    """The roofline `traffic` fields come from these rows; a reformatted
    summary would silently turn them into null."""
    t = bench.pmc_traffic(kernel, profile=profile)
    assert t is not None and t > 0
    assert bench.pmc_traffic("no such kernel", profile=profile) is None
    assert bench.pmc_traffic(kernel, profile="no_such_profile.md") is None


def test_flop_counts():
    """SYRK of an (ov x N) operand: N(N+1)/2 outputs, 2 ov flops each, per
    frequency; Cholesky n^3/3 per frequency."""
    o, v, n, nf = 3, 5, 7, 2
    brute = nf * sum(2 * o * v for p in range(n) for q in range(p + 1))
    assert bench.syrk_flops(o, v, n, nf) == brute
    assert bench.chol_flops(6, 4) == pytest.approx(4 * 216 / 3.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
