"""FP32 precision form (FP32 stepping, FP64 islands, compensated sums) against
the FP64 CPU oracle.

Bar, stated per test: mean relative intensity error over pixels with
I_ref > 1e-6 max(I_ref) <= 1e-3; hit class identical outside the
epsilon band: pixels within 0.1 M of the critical curve (where
classification is chaotic), and rays with a recorded equatorial crossing
within a relative 1e-4 of the disk's inner or outer edge in either image
(where the crossing radius's rounding decides between a hit and a pass).
"""
from __future__ import annotations

import numpy as np
import pytest

from blackhole_amd import abi, analysis, tracer, workloads
from parity_util import run_gpu, run_oracle

pytestmark = pytest.mark.gpu

EPS_BAND = 0.1      # M
EDGE_REL = 1e-4
TOL_MEAN_I = 1e-3


def band_split(cam, opts, diff_idx, got_diff, ref_diff) -> dict:
    """Class disagreements (pixel indices diff_idx within the camera, their
    records in both images) split into the epsilon band and the rest."""
    al, be = analysis.pixel_alpha_beta(cam, diff_idx)
    curve = analysis.curve_distance_points(cam, al, be) <= EPS_BAND
    edge = np.zeros(len(diff_idx), dtype=bool)
    if opts.disk_enabled:
        r_in = tracer.isco(cam.spin)
        for r in (got_diff, ref_diff):
            edge |= analysis.near_disk_edge(r["cross_r"], r["n_cross"], r_in, opts.disk_r_out, EDGE_REL)
    return {"class_diff": int(len(diff_idx)), "in_curve_band": int(curve.sum()),
            "at_disk_edge": int((edge & ~curve).sum()), "class_diff_outside_band": int((~curve & ~edge).sum())}


def _compare(w, oracle, label):
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    cam = w["cams"][0]
    diff = np.nonzero(gpu["hit_class"] != ref["hit_class"])[0]
    same = gpu["hit_class"] == ref["hit_class"]
    err = analysis.intensity_errors(gpu["intensity"][same], ref["intensity"][same])
    stats = dict(err, **band_split(cam, w["opts"], diff, gpu[diff], ref[diff]),
                 island_rays=int((gpu["fp64_steps"] > 0).sum()),
                 island_step_frac=float(gpu["fp64_steps"].sum() / max(gpu["steps"].sum(), 1)))
    print(label, stats, flush=True)
    return gpu, ref, stats


def test_fp32_rk4_disk_c1(oracle):
    w = workloads.c1(192, 192)
    w["opts"].precision = abi.PRECISION_FP32
    gpu, ref, s = _compare(w, oracle, "fp32-c1")
    assert s["class_diff_outside_band"] == 0
    assert s["mean_rel"] <= TOL_MEAN_I
    assert s["island_rays"] > 0
    assert ((gpu["flags"] & abi.FLAG_FP64_ISLAND) != 0).sum() == s["island_rays"]


def test_fp32_far_camera_traces_in_before_escaping(oracle):
    """A camera beyond r_escape (r_obs 600 M against 200 M): the FP32 form's
    rays travel in and leave through r_escape on the way out, as the
    oracle's do."""
    cams = [tracer.camera(0.5, 40.0, 600.0, 96, 96, 15.0)]
    tracer.pack_cameras(cams)
    o = workloads.c5_options()
    o.precision = abi.PRECISION_FP32
    w = dict(cams=cams, opts=o, luts=tracer.disk_luts(cams, 20.0))
    gpu, ref, s = _compare(w, oracle, "fp32-far")
    assert s["class_diff_outside_band"] == 0
    assert s["mean_rel"] <= TOL_MEAN_I
    assert (ref["hit_class"] == abi.HIT_DISK).sum() > 4000 and (ref["hit_class"] == abi.HIT_HORIZON).sum() > 300
    esc = gpu["hit_class"] == abi.HIT_ESCAPED
    assert esc.sum() > 200 and (gpu["r_exit"][esc] > 200.0).all() and (gpu["r_exit"][esc] < 260.0).all()


def test_fp32_dp45_volume_c2(oracle):
    w = workloads.c2(128, 128)
    w["opts"].precision = abi.PRECISION_FP32
    w["opts"].rtol = 1e-6
    w["opts"].atol = 1e-8
    gpu, ref, s = _compare(w, oracle, "fp32-c2")
    assert s["class_diff_outside_band"] == 0
    assert s["mean_rel"] <= TOL_MEAN_I


def test_fp32_precision_form_c3_small(oracle):
    w = workloads.c3(160, 160)
    gpu, ref, s = _compare(w, oracle, "fp32-c3")
    assert s["class_diff_outside_band"] == 0
    assert s["mean_rel"] <= TOL_MEAN_I


def test_fp32_precision_form_c3_full_4096_against_fp64_frame():
    """The full c3 frame (4096^2, a = 0.998, 85 deg, disk + torus) in the FP32
    form against the same frame traced by the FP64 path (which equals the
    oracle bit for bit, tests/test_gpu_parity.py): mean relative intensity
    <= 1e-3 and no class difference outside the epsilon band.  Compared on
    the device; only the disagreeing records come to the host."""
    import torch
    w = workloads.c3(4096, 4096)
    recs = {}
    for prec in (abi.PRECISION_FP32, abi.PRECISION_FP64):
        w["opts"].precision = prec
        print("[c3-4096] tracing precision", prec, flush=True)
        sc = tracer.upload_scene(w["cams"], w["opts"], luts=w["luts"], cells=w["cells"], vol=w["vol"])
        out = tracer.alloc_records(sc.n_records)
        out.fill_(0)
        tracer.launch(sc, out)
        torch.cuda.synchronize()
        recs[prec] = out
    a, b = recs[abi.PRECISION_FP32], recs[abi.PRECISION_FP64]
    assert bool((b[:, 112] != 0).all()) and bool((a[:, 112] != 0).all())
    same = a[:, 112] == b[:, 112]
    Ia, Ib = a.view(torch.float64)[:, 0], b.view(torch.float64)[:, 0]
    m = same & (Ib > 1e-6 * Ib.max())
    rel = (Ia[m] - Ib[m]).abs() / Ib[m]
    diff = torch.nonzero(~same).flatten()
    got = tracer.records_to_numpy(a[diff])
    ref = tracer.records_to_numpy(b[diff])
    w["opts"].precision = abi.PRECISION_FP32
    s = dict(band_split(w["cams"][0], w["opts"], diff.cpu().numpy(), got, ref),
             mean_rel=float(rel.mean()), max_rel=float(rel.max()), n_pixels=int(m.sum()),
             island_step_frac=float(a.view(torch.int32)[:, 30].double().sum() / a.view(torch.int32)[:, 26].double().sum()))
    print("fp32-c3-4096 vs fp64", s, flush=True)
    assert s["mean_rel"] <= TOL_MEAN_I
    assert s["class_diff_outside_band"] == 0
