"""Render-launch ABI on the GPU: per-pixel records of the pinhole camera
against the oracle (FP64 variant), the FP32 variant's hit classes, and the
shaded framebuffer against the records."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as OL
from parity_util import compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spin,pos", [(0.9, (0.0, -20.0, 12.0)), (0.3, (14.0, 6.0, -3.0))])
def test_launch_records_match_oracle(spin, pos):
    import torch
    from blackhole_amd import tracer
    p = tracer.launch_params(rs=2.0, spin=spin, cam_pos=pos, width=48, height=32, fov_scale=0.4)
    out = tracer.trace_launch_params(p, variant=2)
    torch.cuda.synchronize()
    gpu = tracer.records_to_numpy(out)
    ref = OL.trace_launch_params(p)
    st = compare(gpu, ref)
    print("render records", spin, pos, st, np.bincount(gpu["hit_class"], minlength=6), flush=True)
    assert st["class_mismatch"] == 0 and st["bitwise_exit"] == 1.0
    assert st["max_rel_intensity"] <= 1e-9
    assert (gpu["hit_class"] == 2).sum() > 0 and (gpu["hit_class"] == 1).sum() > 0


def test_camera_beyond_max_dist_renders_the_hole():
    """max_dist is the escape radius for outbound rays: a camera at 360 M with
    max_dist = 200 M sees the disk and the shadow (records equal to the
    oracle's), its escaped rays leave through max_dist on the way out."""
    import torch
    from blackhole_amd import tracer
    p = tracer.launch_params(rs=2.0, spin=0.5, cam_pos=(0.0, -300.0, 200.0), width=48, height=32,
                             fov_scale=0.03, max_dist=200.0)
    gpu = tracer.records_to_numpy(tracer.trace_launch_params(p, variant=2))
    torch.cuda.synchronize()
    ref = OL.trace_launch_params(p)
    st = compare(gpu, ref)
    print("far camera render records", st, np.bincount(gpu["hit_class"], minlength=6), flush=True)
    assert st["class_mismatch"] == 0 and st["bitwise_exit"] == 1.0
    assert st["max_rel_intensity"] <= 1e-9
    assert (gpu["hit_class"] == 2).sum() > 500 and (gpu["hit_class"] == 1).sum() > 50
    esc = gpu["hit_class"] == 3
    assert (gpu["r_exit"][esc] > 200.0).all()


def test_fp32_variant_classes_and_framebuffer_shading():
    import torch
    from blackhole_amd import abi, tracer
    p = tracer.launch_params(rs=2.0, spin=0.9, cam_pos=(0.0, -20.0, 12.0), width=64, height=48, fov_scale=0.4,
                             background_enabled=1, background_intensity=0.25)
    rec64 = tracer.records_to_numpy(tracer.trace_launch_params(p, variant=2))
    rec32 = tracer.records_to_numpy(tracer.trace_launch_params(p, variant=0))
    # hit classes identical outside the epsilon band: rays whose constants lie
    # within 0.1 M of the critical curve (in the distant-observer image plane
    # of the camera's inclination) or with a crossing within 1e-4 of a disk edge
    import math
    from blackhole_amd import analysis
    diff = np.nonzero(rec64["hit_class"] != rec32["hit_class"])[0]
    cam_r = math.sqrt(sum(x * x for x in (0.0, -20.0, 12.0)))
    th = math.acos(12.0 / cam_r)
    outside = 0
    for i in diff:
        _, xi, eta = OL.pinhole_setup(p, int(i % p.width), int(i // p.width))
        al = -xi / math.sin(th)
        be = math.sqrt(max(eta + 0.81 * math.cos(th) ** 2 - xi * xi / math.tan(th) ** 2, 0.0))
        near_curve = analysis.curve_distance_points(tracer.camera(0.9, math.degrees(th), cam_r, 1, 1),
                                                    np.array([al]), np.array([be]))[0] <= 0.1
        edge = any(analysis.near_disk_edge(r["cross_r"][i], r["n_cross"][i], tracer.isco(0.9), 60.0, 1e-4)[0]
                   for r in (rec64, rec32))
        outside += int(not (near_curve or edge))
    print("fp32 variant class differences", len(diff), "outside the band", outside, flush=True)
    assert outside == 0
    fb = tracer.render_launch(p, variant=2)
    torch.cuda.synchronize()
    fb = fb.cpu().numpy().reshape(-1, 4)
    cls = rec64["hit_class"]
    assert (fb[cls == abi.HIT_HORIZON] == np.array([0, 0, 0, 1], np.float32)).all()
    disk = cls == abi.HIT_DISK
    assert (fb[disk, 3] == 1.0).all() and (fb[disk, 0] >= 0.0).all() and fb[disk, 0].max() > 0.0
    esc = cls == abi.HIT_ESCAPED
    assert np.allclose(fb[esc], np.array([0.25, 0.25, 0.25, 0.0], np.float32))
