"""Render-launch ABI (include/bh_kernel_launch.h) on the CPU: the pinhole
camera's constants of motion pinned by the distant-observer limit (Bardeen's
image plane) and by the Schwarzschild closed form, the parameter layout
against a compile of the header and the reference's golden offsets, and
argument validation that needs no GPU."""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL
from blackhole_amd import abi, tracer
from blackhole_amd._lib import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dir_of_pixel(p, px, py):
    """The pixel direction exactly as the render ABI maps pixels."""
    u = (2.0 * (px + 0.5) / p.width - 1.0) * p.fov_scale * (p.width / p.height) + p.frame_shift_x
    v = (2.0 * (py + 0.5) / p.height - 1.0) * p.fov_scale + p.frame_shift_y
    loc = np.array([-u, -v, 1.0])
    loc /= np.linalg.norm(loc)
    B = np.array(p.cam_basis[:], dtype=np.float64).reshape(3, 3).T  # columns right, up, forward
    return B @ loc


def _frame(pos):
    r = np.linalg.norm(pos)
    th = math.acos(pos[2] / r)
    ph = math.atan2(pos[1], pos[0])
    e_r = pos / r
    e_t = np.array([math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th)])
    e_p = np.array([-math.sin(ph), math.cos(ph), 0.0])
    return r, th, e_r, e_t, e_p


def test_pinhole_constants_reduce_to_bardeen_image_plane_far_away():
    """A camera at 1e6 M: alpha = -r (d . e_phi), beta = -r (d . e_theta) and
    xi = -alpha sin(theta), eta = beta^2 + (alpha^2 - a^2) cos^2(theta),
    v_r < 0, sign(v_u) = sign(beta) (O(M/r) corrections)."""
    a = 0.7
    for th_deg, ph in ((60.0, 0.3), (25.0, -1.1), (100.0, 2.0)):
        th = math.radians(th_deg)
        R = 1.0e6
        pos = R * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        p = tracer.launch_params(rs=2.0, spin=a, cam_pos=tuple(pos), fov_scale=2e-5, width=9, height=7)
        r, th_c, e_r, e_t, e_p = _frame(np.array(p.cam_pos[:], dtype=np.float64))
        for px, py in ((0, 0), (8, 6), (4, 3), (2, 5)):
            y, xi, eta = OL.pinhole_setup(p, px, py)
            d = _dir_of_pixel(p, px, py)
            alpha = -r * float(d @ e_p)
            beta = -r * float(d @ e_t)
            xi_b = -alpha * math.sin(th_c)
            eta_b = beta * beta + (alpha * alpha - a * a) * math.cos(th_c) ** 2
            scale = 1.0 + alpha * alpha + beta * beta
            assert abs(xi - xi_b) <= 1e-4 * scale, (th_deg, px, py, xi, xi_b)
            assert abs(eta - eta_b) <= 1e-4 * scale, (th_deg, px, py, eta, eta_b)
            assert y[3] < 0.0
            if abs(beta) > 1e-3:
                assert np.sign(y[4]) == np.sign(beta)


def test_pinhole_constants_match_schwarzschild_closed_form():
    """a = 0: with c = d . e_r and s^2 = |d|^2 - c^2 the tangential part of the
    pixel direction d (|d| = 1 to the float32 orthonormality of the camera
    basis), the photon's impact parameter is
    b^2 = eta + xi^2 = r^2 s^2 / (c^2 + (1 - 2/r) s^2)."""
    for pos in ((0.0, -9.0, 4.0), (5.0, 3.0, -2.5), (30.0, 1.0, 0.5)):
        p = tracer.launch_params(rs=2.0, spin=0.0, cam_pos=pos, fov_scale=0.6, width=11, height=9)
        r, th, e_r, e_t, e_p = _frame(np.array(p.cam_pos[:], dtype=np.float64))
        for px, py in ((0, 0), (10, 8), (5, 4), (7, 1)):
            y, xi, eta = OL.pinhole_setup(p, px, py)
            d = _dir_of_pixel(p, px, py)
            c = float(d @ e_r)
            s2 = float(d @ d) - c * c
            b2 = r * r * s2 / (c * c + (1.0 - 2.0 / r) * s2)
            assert abs((eta + xi * xi) - b2) <= 1e-9 * max(b2, 1.0), (pos, px, py, eta + xi * xi, b2)
            # the initial state satisfies the potentials exactly
            assert abs(y[0] - r) <= 1e-12 * r and abs(y[1] - math.cos(th)) <= 1e-12
            assert np.sign(y[3]) == (1.0 if c >= 0 else -1.0)


def test_launch_params_layout_matches_header_and_reference(tmp_path):
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "bh_launch_params_layout.json")))
    fields = [f[0] for f in abi.BH_LaunchParams._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bh_kernel_launch.h"', "int main(void) {",
             '  printf("__sizeof__ %zu\\n", sizeof(struct BH_LaunchParams));']
    for f in fields:
        lines.append(f'  printf("{f} %zu\\n", offsetof(struct BH_LaunchParams, {f}));')
    for e in ("BH_KERNEL_FP32_BASELINE", "BH_KERNEL_FP32_COARSENED", "BH_KERNEL_FP16_STORAGE",
              "BH_KERNEL_FP16_H2_ILP", "BH_KERNEL_COUNT"):
        lines.append(f'  printf("{e} %d\\n", (int){e});')
    lines += ["  return 0;", "}"]
    src = tmp_path / "lp.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "lp"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.strip().split("\n"))
    for k, v in got.items():
        assert int(v) == golden[k], (k, v, golden[k])
        if k in fields:
            assert int(v) == getattr(abi.BH_LaunchParams, k).offset


@pytest.mark.parametrize("field,value,msg", [
    ("cam_pos", (0.0, 0.0, 30.0), b"spin axis"),
    ("cam_pos", (0.5, 0.5, 0.2), b"outside r = rs"),
    ("spin", 1.0, b"spin"),
    ("width", 0, b"empty"),
    ("max_steps", 0, b"max_steps"),
])
def test_render_launch_rejects_invalid_params_without_gpu(field, value, msg):
    p = tracer.launch_params()
    if field == "cam_pos":
        for i in range(3):
            p.cam_pos[i] = value[i]
    else:
        setattr(p, field, value)
    assert lib().bh_trace_launch_params(C.byref(p), 2, C.c_void_p(1), None) == -1
    assert msg in lib().bh_last_error()
    assert lib().bh_launch_geodesic_kernel(None, C.byref(p), 2, None) == -1


def test_texture_handles_and_variant_selection():
    lib().bh_upload_lut_textures(1, 2, 3, 4, 5, 6, 7, 8)
    lib().bh_upload_bridge_background_texture(99)
    assert [lib().bh_lut_texture_handle(i) for i in range(9)] == [1, 2, 3, 4, 5, 6, 7, 8, 99]
    assert lib().bh_lut_texture_handle(9) == 0 and lib().bh_lut_texture_handle(-1) == 0
    lib().bh_upload_lut_textures(0, 0, 0, 0, 0, 0, 0, 0)
    assert lib().bh_select_kernel_variant() == 2


def test_bridge_probes_match_reference_struct_sizes():
    """sizeof probes: BhbSourceParams = char[64] + 7 doubles = 120 bytes and
    BhbDiskParams = 5 doubles = 40 bytes (blender_bridge.h:48-57, :180-186);
    versions 1.0; snake_case aliases exported."""
    L = lib()
    assert L.bhbSizeofSourceParams() == 120 and L.bhbSizeofDiskParams() == 40
    assert L.bhb_sizeof_source_params() == 120 and L.bhb_sizeof_disk_params() == 40
    assert (L.bhbVersionMajor(), L.bhbVersionMinor(), L.bhbHasCuda(), L.bhbHasBoost()) == (1, 0, 1, 0)
    assert (L.bhb_version_major(), L.bhb_has_cuda()) == (1, 1)


def test_bridge_rejects_invalid_arguments_without_gpu():
    L = lib()
    f = L.bhbCudaTraceGeodesics
    f.argtypes = [C.c_double] * 5 + [C.c_int] + [C.c_double] * 2 + [C.c_int, C.c_int, C.c_double, C.c_void_p,
                                                                   C.c_void_p]
    buf = (C.c_float * 300)()
    cnt = (C.c_int * 4)()
    assert f(1.0, 50.0, 1.0, -5, 5, 2, -5, 5, 2, 10, 0.1, buf, cnt) == -1 and b"spin" in L.bh_last_error()
    assert f(0.5, 1.1, 1.0, -5, 5, 2, -5, 5, 2, 10, 0.1, buf, cnt) == -1 and b"horizon" in L.bh_last_error()
    assert f(0.5, 50.0, 0.0, -5, 5, 2, -5, 5, 2, 10, 0.1, buf, cnt) == -1 and b"inclination" in L.bh_last_error()
    assert f(0.5, 50.0, 1.0, -5, 5, 2, -5, 5, 2, 10, 0.1, None, cnt) == -1
    g = L.bhbCudaRenderLensingMap
    g.argtypes = [C.c_double] * 3 + [C.c_int, C.c_int, C.c_void_p]
    assert g(0.5, 50.0, 1.0, 0, 4, buf) == -1


def test_backend_handles_null_and_empty_without_gpu():
    L = lib()
    assert L.bhCudaInit(0, 0, 16) is None
    assert "empty framebuffer" in L.bh_last_error().decode()
    assert L.bhCudaInit(0, 16, -1) is None
    L.bhCudaShutdown(None)
    assert L.bhCudaGetVariant(None) == -1 and L.bhCudaSetVariant(None, 0) == -1
    assert L.bhCudaRenderFrame(None, None) == -1
    assert L.bhCudaResize(None, 0, 4, 4) == -1
    assert L.bhCudaRegisterLut(None, 0, 1, 0) == -1
    assert L.bhCudaSynchronize(None) == -1
    assert L.bhCudaFramebuffer(None, None, None) is None
    assert L.bhCudaGlTexture(None, -1) == 0
