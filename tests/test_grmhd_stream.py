"""GRMHD frame cache logic on the host (blackhole_amd/csrc/grmhd_stream.hpp,
compiled by g++ with a host-memory backend, tests/cpp/grmhd_stream_check.cpp):
miss / hit semantics, seek prefetch, least-recently-used eviction under the
budget with the current and next frames kept, failed loads, shutdown
releasing every frame.  The device backend is tested in test_gpu_grmhd_stream.py."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_frame_cache_semantics(tmp_path):
    exe = tmp_path / "grmhd_stream_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Wextra",
                    os.path.join(ROOT, "tests", "cpp", "grmhd_stream_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
