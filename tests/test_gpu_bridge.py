"""Bridge entry points (include/bh_bridge.h) on the GPU: image-plane
polylines against the oracle's trajectories, the lensing map against the
oracle's disk crossings, and the ray-traced render's shadow."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as OL
from blackhole_amd import abi, tracer
from blackhole_amd._lib import lib

pytestmark = pytest.mark.gpu


def _bridge_camera(a, r_obs, inc, amin, amax, na, bmin, bmax, nb):
    """The BH_Camera whose pixel centres are the bridge's inclusive grid."""
    da = (amax - amin) / (na - 1)
    db = (bmax - bmin) / (nb - 1)
    cam = abi.BH_Camera(a, inc, r_obs, amin - 0.5 * da, amin - 0.5 * da + na * da,
                        bmin - 0.5 * db + nb * db, bmin - 0.5 * db, na, nb, 0)
    return cam


def test_polylines_end_at_the_oracle_exit_state():
    import torch  # noqa: F401  (loads the HIP runtime first)
    a, r_obs, inc = 0.8, 60.0, math.radians(70.0)
    na, nb, ms = 9, 7, 400
    xyz = np.zeros(na * nb * ms * 3, np.float32)
    cnt = np.zeros(na * nb, np.int32)
    L = lib()
    L.bhbCudaTraceGeodesics.argtypes = [C.c_double] * 5 + [C.c_int] + [C.c_double] * 2 + [C.c_int, C.c_int, C.c_double,
                                                                                          C.c_void_p, C.c_void_p]
    rc = L.bhbCudaTraceGeodesics(a, r_obs, inc, -12.0, 12.0, na, -9.0, 9.0, nb, ms, 0.05,
                                 xyz.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p))
    assert rc == 0, L.bh_last_error()
    cam = _bridge_camera(a, r_obs, inc, -12.0, 12.0, na, -9.0, 9.0, nb)
    opts = tracer.default_options(integrator=abi.INTEGRATOR_DP45, disk_enabled=0, max_crossings=3,
                                  max_steps=ms - 1, max_attempts=20 * ms + 1000, r_escape=3.0 * r_obs)
    ref = OL.trace([cam], opts)
    pts = xyz.reshape(na * nb, ms, 3)
    assert (cnt == np.minimum(ref["steps"] + 1, ms)).all()
    first = pts[:, 0, :]
    assert np.allclose(first, [r_obs * math.sin(inc), 0.0, r_obs * math.cos(inc)], rtol=1e-6)
    for i in range(na * nb):
        r, u, ph = ref["r_exit"][i], ref["u_exit"][i], ref["phi_exit"][i]
        st = math.sqrt(max(1.0 - u * u, 0.0))
        want = np.array([r * st * math.cos(ph), r * st * math.sin(ph), r * u], np.float32)
        assert np.allclose(pts[i, cnt[i] - 1], want, rtol=1e-6, atol=1e-5), (i, pts[i, cnt[i] - 1], want)
    # the reference's grid order: beta rows from betaMin up, alpha columns from alphaMin
    assert set(np.unique(ref["hit_class"])) >= {abi.HIT_HORIZON, abi.HIT_ESCAPED}


def test_lensing_map_matches_oracle_crossings():
    import torch  # noqa: F401
    a, r_obs, inc, w, h = 0.6, 50.0, math.radians(60.0), 40, 30
    out = np.zeros(w * h * 4, np.float32)
    L = lib()
    L.bhbCudaRenderLensingMap.argtypes = [C.c_double] * 3 + [C.c_int, C.c_int, C.c_void_p]
    assert L.bhbCudaRenderLensingMap(a, r_obs, inc, w, h, out.ctypes.data_as(C.c_void_p)) == 0, L.bh_last_error()
    da, db = 30.0 / w, 30.0 / h
    cam = abi.BH_Camera(a, inc, r_obs, -15.0 - 0.5 * da, -15.0 - 0.5 * da + w * da,
                        15.0 + 0.5 * db - h * db, 15.0 + 0.5 * db, w, h, 0)
    opts = tracer.default_options(integrator=abi.INTEGRATOR_DP45, max_steps=20000, disk_enabled=1,
                                  disk_r_out=100.0, r_escape=3.0 * r_obs)
    r_isco = tracer.isco(a)
    lut = OL.disk_lut(a, r_isco, 100.0)
    ref = OL.trace([cam], opts, luts=lut.reshape(1, -1), lut_n=lut.size)
    px = out.reshape(-1, 4)
    disk = ref["hit_class"] == abi.HIT_DISK
    assert disk.sum() > 0
    assert (px[:, 2] == disk.astype(np.float32)).all()
    k = np.minimum(ref["n_cross"][disk], 3) - 1
    rc = ref["cross_r"][disk][np.arange(k.size), k]
    g = ref["cross_g"][disk][np.arange(k.size), k]
    grav = np.sqrt(1.0 - 2.0 / rc)
    assert np.allclose(px[disk, 0], grav, rtol=1e-6) and np.allclose(px[disk, 1], g / grav, rtol=1e-6)
    cap = ref["hit_class"] == abi.HIT_HORIZON
    assert (px[cap, 0] == 0.0).all() and (px[:, 3] == 1.0).all()


def test_raytraced_render_shows_the_shadow():
    import torch  # noqa: F401
    w, h = 64, 48
    out = np.zeros(w * h * 4, np.float32)
    L = lib()
    L.bhbCudaRenderRaytraced.argtypes = [C.c_float] * 3 + [C.c_int, C.c_int, C.c_void_p]
    # nearly face-on: the disk's inner edge surrounds the shadow instead of crossing in front of it
    assert L.bhbCudaRenderRaytraced(0.9, 30.0, math.radians(20.0), w, h, out.ctypes.data_as(C.c_void_p)) == 0, \
        L.bh_last_error()
    img = out.reshape(h, w, 4)
    centre = img[h // 2 - 1:h // 2 + 1, w // 2 - 1:w // 2 + 1]
    assert (centre[..., :3] == 0.0).all() and (centre[..., 3] == 1.0).all()
    assert (img[..., 3] == 0.0).any() and img[..., 0].max() > 0.0
