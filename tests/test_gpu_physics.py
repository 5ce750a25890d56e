"""The physics C++ interface's GPU tracing (KerrRaytracer over
bh_trace_rays_host) against the oracle: equatorial rays of the interface map
to image-plane rays whose exit states the oracle reproduces."""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL
from blackhole_amd import abi, tracer

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_kerr_raytracer_matches_oracle(tmp_path):
    exe = tmp_path / "physics_api_check"
    lib_dir = os.path.join(ROOT, "blackhole_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "physics_api_check.cpp"), "-o", str(exe),
                    "-L", lib_dir, "-lbhk", f"-Wl,-rpath,{lib_dir}"], check=True)
    print("[physics] raytrace", flush=True)
    out = subprocess.run([str(exe), "raytrace"], capture_output=True, text=True, check=True, timeout=120).stdout
    rows = np.array([[float(v) for v in l.split()] for l in out.strip().split("\n")])
    # status: CAPTURED = 2, ESCAPED = 1; capture iff b below the critical impact parameter
    statuses = rows[:, 0].astype(int)
    assert statuses[0] == 2 and statuses[-1] == 1
    # the same rays as image-plane pixels: alpha = -b (xi = b at theta = pi/2), beta = 0
    opts = tracer.default_options(integrator=abi.INTEGRATOR_DP45, disk_enabled=0, max_steps=20000,
                                  max_attempts=10 * 20000 + 1000, horizon_factor=1.001, r_escape=200.0)
    for (st, steps, r, th, ph), b in zip(rows, (4.0, 5.0, 6.0, 7.0)):
        cam = abi.BH_Camera(0.5, math.pi / 2, 80.0, -b - 0.5, -b + 0.5, -0.5, 0.5, 1, 1, 0)
        ref = OL.trace([cam], opts)
        want = {abi.HIT_HORIZON: 2, abi.HIT_ESCAPED: 1}[int(ref["hit_class"][0])]
        assert int(st) == want
        assert abs(r - ref["r_exit"][0]) <= 1e-8 * ref["r_exit"][0]
        assert abs(ph - ref["phi_exit"][0]) <= 1e-8 * max(1.0, abs(ref["phi_exit"][0]))
        assert abs(steps - ref["steps"][0]) <= 2


def test_backend_abi_renders_into_its_framebuffer():
    import ctypes as C
    import torch
    from blackhole_amd._lib import lib
    L = lib()
    b = L.bhCudaInit(7, 32, 24)
    assert b
    try:
        assert L.bhCudaGetVariant(b) == 2
        assert L.bhCudaSetVariant(b, 0) == 0 and L.bhCudaGetVariant(b) == 0
        assert L.bhCudaSetVariant(b, 99) == 0  # beyond the table -> FP32 baseline
        assert L.bhCudaSetVariant(b, -1) == 2  # negative -> the selected default
        assert L.bhCudaRegisterLut(b, 3, 11, 0) == 0 and L.bhCudaRegisterBackgroundTexture(b, 12, 0) == 0
        assert L.bhCudaGlTexture(b, -1) == 7 and L.bhCudaGlTexture(b, 3) == 11 and L.bhCudaGlTexture(b, 8) == 12
        p = tracer.launch_params(width=32, height=24)
        print("[backend] render 32x24", flush=True)
        assert L.bhCudaRenderFrame(b, C.byref(p)) == 0 and L.bhCudaSynchronize(b) == 0
        w, h = C.c_int(), C.c_int()
        ptr = L.bhCudaFramebuffer(b, C.byref(w), C.byref(h))
        assert ptr and (w.value, h.value) == (32, 24)
        # the backend's framebuffer equals a direct launch of the same params and variant
        got = torch.empty((24, 32, 4), dtype=torch.float32, device="cuda")
        # the HIP runtime torch and libbhk.so already share (RTLD_NOLOAD: never map a second copy)
        import os
        hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_LAZY)
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        assert hip.hipMemcpy(C.c_void_p(got.data_ptr()), C.c_void_p(ptr), got.numel() * 4, 3) == 0  # D2D
        ref = tracer.render_launch(p, variant=2)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        assert L.bhCudaResize(b, 9, 16, 8) == 0 and L.bhCudaGlTexture(b, -1) == 9
        assert L.bhCudaRenderFrame(b, C.byref(p)) == -1  # size differs from the framebuffer
        p2 = tracer.launch_params(width=16, height=8)
        assert L.bhCudaRenderFrame(b, C.byref(p2)) == 0 and L.bhCudaSynchronize(b) == 0
    finally:
        L.bhCudaShutdown(b)
