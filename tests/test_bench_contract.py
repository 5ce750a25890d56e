"""bench.py's contract on the CPU: the refusal of more than one GPU before
any device work, a loud failure without a device, and the measured-peak
fields that every roofline of the JSON line carries."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=300, cwd=ROOT)


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_is_refused_with_a_message(n):
    r = _run("--gpus", n)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_without_a_device_the_bench_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    r = _run("--steps", "1", "--warmup", "0")
    assert r.returncode == 1 and "no HIP device" in r.stderr and r.stdout == ""


def test_measured_peaks_attach_to_every_roofline():
    sys.path.insert(0, ROOT)
    import bench
    pk = {"fp64_fma_tflops": 64.0, "fp32_fma_scalar_tflops": 70.0, "bf16_mfma_tflops": 2000.0,
          "hbm_copy_gbs": 5000.0, "note": ""}
    res = {"roofline": {"bound": "mfma", "achieved": 1000.0, "peak": bench.BF16_DENSE_PEAK_TFLOPS},
           "plain_dp45": {"roofline": {"bound": "valu", "achieved": 32.0, "peak": bench.FP64_VECTOR_PEAK_TFLOPS}},
           "legs": [{"roofline": {"achieved": 35.0, "peak": 157.3}}],
           "other": {"achieved": 1.0, "peak": 8.0}}  # an HBM figure: no measured twin
    bench.attach_measured_peaks(res, pk)
    assert res["roofline"]["peak_measured"] == 2000.0 and res["roofline"]["frac_measured"] == 0.5
    assert res["plain_dp45"]["roofline"]["frac_measured"] == 0.5
    assert res["legs"][0]["roofline"]["frac_measured"] == 0.5
    assert "peak_measured" not in res["other"]
    assert res["measured_peaks"] is pk


def test_product_path_fails_loudly_without_the_library():
    """No CPU fallback: with libbhk.so absent, the first call into the
    product raises NativeLibraryMissing (run in a child process so the
    loaded library of this one is untouched)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from blackhole_amd import _lib, tracer\n"
        "_lib.LIB_PATH = '/nonexistent/libbhk.so'\n"
        "try:\n"
        "    tracer.outer_horizon(0.5)\n"
        "except _lib.NativeLibraryMissing as e:\n"
        "    print('raised:', e)\n"
        "    sys.exit(0)\n"
        "sys.exit(3)\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "is not built" in r.stdout
