"""The escape test r > r_escape applies to outbound rays only.

For a camera inside r_escape this is the plain r > r_escape (an inbound ray
reaches r_escape only after its radial turning point); a camera beyond it
(c4 draws r_obs up to 1000 M against r_escape = 200 M) has its rays travel
in first.  Pinned without reference to the rule's own code: a far camera's
records with r_escape = 200 M must carry the same physics -- hit class,
crossings, g, intensity -- as with r_escape = 2000 M beyond the camera, where
every ray starts inbound inside the escape sphere and the two forms of the
test coincide; escaped rays end on the way out just past 200 M; and the
Mino-time solver, whose path ends at r_escape on the outbound leg by
construction, agrees with the oracle's crossings.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL
from blackhole_amd import abi, tracer, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _scene(r_obs, r_escape, size=20, spin=0.7, inc=55.0):
    cams = [tracer.camera(spin, inc, r_obs, size, size, 15.0)]
    tracer.pack_cameras(cams)
    o = workloads.c5_options()
    o.r_escape = r_escape
    return cams, o, tracer.disk_luts(cams, 20.0)


def test_far_camera_rays_travel_in_before_they_escape():
    cams, o_near, luts = _scene(500.0, 200.0)
    _, o_far, _ = _scene(500.0, 2000.0)
    a = OL.trace(cams, o_near, luts=luts, lut_n=luts.shape[1], nthreads=8)
    b = OL.trace(cams, o_far, luts=luts, lut_n=luts.shape[1], nthreads=8)
    cls = a["hit_class"]
    assert (cls == b["hit_class"]).all()
    assert (cls == abi.HIT_DISK).sum() > 40 and (cls == abi.HIT_HORIZON).sum() > 20
    assert (cls == abi.HIT_ESCAPED).sum() > 40
    # the same path up to the escape sphere: crossings, g and intensity agree
    # to the step controller's noise (the two runs take different last steps
    # only on the way out, past every crossing)
    for f in ("cross_r", "cross_g"):
        assert np.allclose(a[f], b[f], rtol=1e-9, atol=0.0), f
    assert np.allclose(a["intensity"], b["intensity"], rtol=1e-9, atol=0.0)
    assert (a["n_cross"] == b["n_cross"]).all()
    esc = cls == abi.HIT_ESCAPED
    # escaped rays end on the way out just beyond r_escape (not at r_obs after one step)
    assert (a["r_exit"][esc] > 200.0).all() and (a["r_exit"][esc] < 260.0).all()
    assert (a["steps"][esc] > 20).all()
    assert (b["r_exit"][esc] > 2000.0).all()


def test_camera_inside_r_escape_is_unchanged_by_the_outbound_condition():
    """r_obs = 100 < r_escape: every escaped ray passes r_escape outbound, so
    its exit radius is within one step of r_escape and above it."""
    cams, o, luts = _scene(100.0, 200.0)
    a = OL.trace(cams, o, luts=luts, lut_n=luts.shape[1], nthreads=8)
    esc = a["hit_class"] == abi.HIT_ESCAPED
    assert esc.sum() > 40
    assert (a["r_exit"][esc] > 200.0).all() and (a["r_exit"][esc] < 260.0).all()


@pytest.fixture(scope="module")
def mino(tmp_path_factory):
    so = tmp_path_factory.mktemp("mino_esc") / "libmino_cpu.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    os.path.join(ROOT, "tests", "cpp", "mino_capi.cpp"), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    D = C.c_double
    L.mino_solve_pixel.argtypes = [D] * 9 + [C.POINTER(D), C.c_int, D, C.POINTER(D), C.POINTER(C.c_int)]
    L.mino_solve_pixel_exit.restype = C.c_int
    L.mino_solve_pixel_exit.argtypes = [D] * 10 + [C.POINTER(D), C.POINTER(C.c_int), C.POINTER(D)]
    return L


def test_mino_solver_matches_the_oracle_for_a_far_camera(mino):
    cams, o, luts = _scene(500.0, 200.0, size=16)
    ref = OL.trace(cams, o, luts=luts, lut_n=luts.shape[1], nthreads=8)
    cam = cams[0]
    D = C.c_double
    r, ints, d = (D * 3)(), (C.c_int * 6)(), (D * 2)()
    n = n_esc = 0
    for j in range(cam.height):
        for i in range(cam.width):
            rec = ref[j * cam.width + i]
            al = cam.alpha_min + (i + 0.5) * (cam.alpha_max - cam.alpha_min) / cam.width
            be = cam.beta_max - (j + 0.5) * (cam.beta_max - cam.beta_min) / cam.height
            mino.mino_solve_pixel(cam.spin, cam.theta_obs, cam.r_obs, al, be,
                                  o.horizon_factor * tracer.outer_horizon(cam.spin), o.r_escape,
                                  tracer.isco(cam.spin), o.disk_r_out, None, 0, 1e-10, r, ints)
            if not ints[4]:
                continue
            assert ints[2] == rec["hit_class"], (i, j)
            for c in range(min(3, ints[0], rec["n_cross"])):
                assert abs(r[c] - rec["cross_r"][c]) <= 1e-6 * rec["cross_r"][c], (i, j, c)
            if rec["hit_class"] == abi.HIT_ESCAPED:
                # the direction at the oracle's own exit radius on the way out
                assert mino.mino_solve_pixel_exit(cam.spin, cam.theta_obs, cam.r_obs, al, be,
                                                  o.horizon_factor * tracer.outer_horizon(cam.spin),
                                                  rec["r_exit"], tracer.isco(cam.spin), o.disk_r_out, 1e-10,
                                                  r, ints, d)
                assert abs(d[0] - rec["u_exit"]) <= 1e-6 and abs(d[1] - rec["phi_exit"]) <= 2e-5, (i, j)
                n_esc += 1
            n += 1
    assert n >= 250 and n_esc > 40
