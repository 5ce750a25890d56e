"""FP32 precision form (FP32 stepping, FP64 islands, compensated sums) against
the FP64 CPU oracle.

Bar, stated per test: mean relative intensity error over pixels with
I_ref > 1e-6 max(I_ref) <= 1e-3; hit-class identical outside an epsilon band
of 0.1 M around the critical curve (inside it classification is chaotic).
"""
from __future__ import annotations

import numpy as np
import pytest

from blackhole_amd import abi, analysis, tracer, workloads
from parity_util import run_gpu, run_oracle

pytestmark = pytest.mark.gpu

EPS_BAND = 0.1      # M
TOL_MEAN_I = 1e-3


def _compare(w, oracle, label):
    gpu = run_gpu(w)
    ref = run_oracle(oracle, w)
    cam = w["cams"][0]
    d = analysis.curve_distance(cam).reshape(-1)
    diff = gpu["hit_class"] != ref["hit_class"]
    same = ~diff
    err = analysis.intensity_errors(gpu["intensity"][same], ref["intensity"][same])
    stats = dict(err, class_diff=int(diff.sum()), class_diff_outside_band=int((diff & (d > EPS_BAND)).sum()),
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
    assert s["class_diff_outside_band"] <= 2
    assert s["mean_rel"] <= TOL_MEAN_I
