"""GPU FP64 path vs the FP64 CPU oracle, through the C ABI (bh_trace_image_plane).

Bar (FP64 path): hit class identical for every ray, exit radius / cos(theta) /
phi within 1e-9 relative, intensity and optical depth within 1e-9 relative.
The trajectory arithmetic is +, -, *, /, sqrt and fma only on both sides, in
the same order, so the expected result is bitwise equality; the tolerance
leaves room only for the transcendentals of the radiative part (acos, exp,
expm1, cbrt, log: libm on the host, the device library on the GPU).
"""
from __future__ import annotations

import numpy as np
import pytest

from blackhole_amd import abi, tracer, workloads
from parity_util import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _check(gpu, ref, label):
    s = compare(gpu, ref)
    print(label, s, flush=True)
    assert s["class_mismatch"] == 0, s
    for f in ("r_exit", "u_exit", "phi_exit", "intensity", "tau"):
        assert s[f"max_rel_{f}"] <= TOL, (f, s)
    assert s["steps_mismatch"] == 0 and s["attempts_mismatch"] == 0, s
    return s


def test_c1_rk4_disk_full_frame_128(oracle):
    w = workloads.c1(128, 128)
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    s = _check(gpu, ref, "c1-128")
    assert s["bitwise_exit"] == 1.0
    assert set(np.unique(gpu["hit_class"])) <= {abi.HIT_HORIZON, abi.HIT_DISK, abi.HIT_ESCAPED}
    assert (gpu["hit_class"] == abi.HIT_DISK).sum() > 1000


def test_c1_rk4_full_512_subset(oracle):
    w = workloads.c1(512, 512)
    gpu = run_gpu(w)
    rng = np.random.default_rng(11)
    ids = np.sort(rng.choice(512 * 512, 4096, replace=False))
    ref = run_oracle(oracle, w, ids=ids)
    _check(gpu[ids], ref, "c1-512-subset")
    # every pixel was traced
    assert (gpu["hit_class"] != abi.HIT_NONE).all()


def test_c2_dp45_volume_128(oracle):
    w = workloads.c2(128, 128)
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    s = _check(gpu, ref, "c2-128")
    assert (gpu["tau"] > 0).sum() > 1000
    assert s["max_rel_intensity"] <= TOL


def test_c2_dp45_volume_full_1024_subset(oracle):
    w = workloads.c2(1024, 1024)
    gpu = run_gpu(w)
    rng = np.random.default_rng(12)
    ids = np.sort(rng.choice(1024 * 1024, 2048, replace=False))
    ref = run_oracle(oracle, w, ids=ids)
    _check(gpu[ids], ref, "c2-1024-subset")
    assert (gpu["hit_class"] != abi.HIT_NONE).all()


def test_dp45_disk_no_volume(oracle):
    w = workloads.c1(96, 96)
    w["opts"] = tracer.default_options(
        integrator=abi.INTEGRATOR_DP45, disk_enabled=1, volume_enabled=0, max_steps=20000,
        nu_obs=workloads.KEV_HZ, disk_T_max=1e7)
    _check(run_gpu(w), run_oracle(oracle, w), "dp45-disk")


def test_ragged_multi_camera(oracle):
    cams = [tracer.camera(0.3, 40.0, 60.0, 1, 1, 10.0),
            tracer.camera(0.7, 80.0, 300.0, 7, 3, 12.0),
            tracer.camera(0.998, 5.0, 1000.0, 33, 17, 8.0),
            tracer.camera(0.0, 89.0, 100.0, 64, 2, 15.0)]
    n = tracer.pack_cameras(cams)
    cams[2].record_offset += 5  # a gap between record ranges is allowed
    cams[3].record_offset += 5
    w = dict(cams=cams, opts=tracer.default_options(max_steps=4000), luts=tracer.disk_luts(cams, 20.0))
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    assert gpu.size == n + 5
    valid = ref["hit_class"] != abi.HIT_NONE
    assert valid.sum() == n
    _check(gpu[valid], ref[valid], "ragged")
    assert (gpu["hit_class"][~valid] == abi.HIT_NONE).all()


def test_max_steps_budget(oracle):
    w = workloads.c1(32, 32)
    w["opts"].max_steps = 50
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    _check(gpu, ref, "budget")
    assert (gpu["hit_class"] == abi.HIT_MAX_STEPS).sum() > 0


def test_repeat_launch_identical():
    w = workloads.c2(64, 64)
    a = run_gpu(w)
    b = run_gpu(w)
    assert a.tobytes() == b.tobytes()
