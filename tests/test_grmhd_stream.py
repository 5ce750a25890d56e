"""GRMHD frame cache logic on the host (blackhole_amd/csrc/grmhd_stream.hpp,
compiled by g++ with a host-memory backend, tests/cpp/grmhd_stream_check.cpp):
miss / hit semantics, seek prefetch, least-recently-used eviction under the
budget with the current and next frames kept, failed loads, shutdown
releasing every frame.  The device backend is tested in test_gpu_grmhd_stream.py."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_frame_cache_semantics(tmp_path):
    exe = tmp_path / "grmhd_stream_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Wextra",
                    os.path.join(ROOT, "tests", "cpp", "grmhd_stream_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_frame_cache_under_sanitizers(tmp_path, sanitizer):
    """The same checks with the cache, its loader thread and the host backend
    built under AddressSanitizer + UndefinedBehaviorSanitizer, and under
    ThreadSanitizer (the loader thread against the render thread): no report."""
    exe = tmp_path / "grmhd_stream_check_san"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fno-omit-frame-pointer",
                    f"-fsanitize={sanitizer}", "-fno-sanitize-recover=all",
                    os.path.join(ROOT, "tests", "cpp", "grmhd_stream_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok" and "Sanitizer" not in r.stderr, r.stdout + r.stderr
