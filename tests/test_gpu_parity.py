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
    # the conserved-quantity drift of each record (null condition v_r^2 = R
    # and the polar first integral v_u^2 = U at the exit state) is computed
    # from the same final state on both sides
    assert s["max_abs_drift_null_diff"] <= 1e-12 and s["max_abs_drift_carter_diff"] <= 1e-12, s
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


def test_pixel_id_list_matches_full_frame():
    import torch
    w = workloads.c1(96, 96)
    full = run_gpu(w)
    sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"])
    rng = np.random.default_rng(21)
    ids = np.sort(rng.choice(96 * 96, 777, replace=False)).astype(np.int64)
    out = tracer.alloc_records(sc.n_records)
    out.zero_()
    d_ids = torch.from_numpy(ids).cuda()
    cnt = torch.tensor([ids.size], dtype=torch.int64, device="cuda")
    tracer.launch_ids(sc, d_ids, cnt, out)
    torch.cuda.synchronize()
    rec = tracer.records_to_numpy(out)
    assert rec[ids].tobytes() == full[ids].tobytes()
    mask = np.ones(96 * 96, bool)
    mask[ids] = False
    assert (rec["hit_class"][mask] == 0).all()
    # a device count smaller than the list traces only its prefix
    out.zero_()
    cnt.fill_(100)
    tracer.launch_ids(sc, d_ids, cnt, out)
    torch.cuda.synchronize()
    rec = tracer.records_to_numpy(out)
    assert (rec["hit_class"][ids[:100]] > 0).all() and (rec["hit_class"][ids[100:]] == 0).all()


def test_explicit_rays_match_camera_rays():
    """Ray-list tracing (device sin/cos/ISCO) agrees with the camera path up to
    the last-ulp difference of the inclination's sin/cos."""
    import torch
    w = workloads.c1(64, 64)
    w["opts"].disk_enabled = 0
    cam = w["cams"][0]
    full = run_gpu(w)
    from blackhole_amd import analysis
    A, B = analysis.pixel_grid(cam)
    specs = np.stack([A.reshape(-1), B.reshape(-1), np.full(A.size, cam.spin),
                      np.full(A.size, cam.theta_obs), np.full(A.size, cam.r_obs)], axis=1)
    out = tracer.trace_rays(torch.from_numpy(specs).cuda(), w["opts"])
    torch.cuda.synchronize()
    rec = tracer.records_to_numpy(out)
    d = analysis.curve_distance(cam).reshape(-1)
    far = d > 0.1
    assert (rec["hit_class"][far] == full["hit_class"][far]).all()
    rel = np.abs(rec["r_exit"][far] - full["r_exit"][far]) / full["r_exit"][far]
    assert np.median(rel) < 1e-12 and rel.max() < 1e-6
    same = far & (full["n_cross"] > 0)
    assert np.abs(rec["cross_r"][same, 0] - full["cross_r"][same, 0]).max() < 1e-6
