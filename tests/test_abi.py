"""C-ABI checks that need no GPU: the library loads, exports every declared
symbol, and its host-side entry points and argument validation behave."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from blackhole_amd import abi, tracer
from blackhole_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(bhb?_?\w+|bh\w+)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name.startswith(("bh", "bhb")) and not name.isupper():
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = declared_functions()
    assert "bh_trace_image_plane" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_defaults():
    assert lib().bh_version() == 1
    o = tracer.default_options()
    assert o.integrator == abi.INTEGRATOR_RK4 and o.max_steps == 2000
    assert o.horizon_factor == 1.01 and o.r_escape == 200.0 and o.disk_r_out == 20.0


def test_closed_forms_match_oracle(oracle):
    for a in (0.0, 0.5, 0.9, 0.94, 0.998):
        assert tracer.isco(a) == oracle.lib().orc_isco(a)
        assert tracer.outer_horizon(a) == oracle.lib().orc_outer_horizon(a)


def test_disk_lut_matches_oracle_bitwise(oracle):
    for a in (0.0, 0.9, 0.998):
        ri = tracer.isco(a)
        cam = tracer.camera(a, 60.0, 100.0, 4, 4)
        got = tracer.disk_luts([cam], 20.0)[0]
        ref = oracle.disk_lut(a, ri, 20.0, 256)
        assert got.tobytes() == ref.tobytes()


def test_torus_k2_lane_matches_oracle_bessel(oracle):
    """The generator's fourth cell lane is K_2(1/Theta_e)/(2 Theta_e^2) of the
    cell's own Theta_e (the oracle's quadrature; pinned against scipy kv)."""
    p = tracer.torus_params(0.94, n_r=32, n_th=16, n_ph=8)
    cells, vol = tracer.generate_torus(p)
    flat = cells.reshape(-1, 4)
    live = flat[flat[:, 1] > 0]
    assert live.shape[0] > 10
    for th, k2r in live[:: max(1, live.shape[0] // 40), 1:4:2]:
        ref = oracle.lib().orc_bessel_k(2.0, 1.0 / float(th)) / (2.0 * float(th) ** 2)
        assert abs(float(k2r) - ref) <= 1e-6 * ref


@pytest.mark.parametrize("field,value", [
    ("integrator", 7), ("max_crossings", 4), ("h_fixed", 0.0), ("precision", 9)])
def test_invalid_options_rejected_without_gpu(field, value):
    o = tracer.default_options()
    setattr(o, field, value)
    cams = (abi.BH_Camera * 1)(tracer.camera(0.9, 60.0, 100.0, 8, 8))
    rc = lib().bh_trace_image_plane(cams, 1, C.byref(o), C.c_void_p(1), 256, None,
                                    C.c_void_p(1), None)
    assert rc == -1
    assert lib().bh_last_error()


@pytest.mark.parametrize("spin,theta,w", [(1.2, 1.0, 8), (0.5, 0.0, 8), (0.5, 1.0, 0)])
def test_invalid_camera_rejected_without_gpu(spin, theta, w):
    o = tracer.default_options()
    cams = (abi.BH_Camera * 1)(abi.BH_Camera(spin, theta, 100.0, -1, 1, -1, 1, w, 8, 0))
    rc = lib().bh_trace_image_plane(cams, 1, C.byref(o), C.c_void_p(1), 256, None,
                                    C.c_void_p(1), None)
    assert rc == -1


def test_null_arguments_rejected():
    o = tracer.default_options()
    assert lib().bh_trace_image_plane(None, 0, C.byref(o), None, 0, None, None, None) == -1
    cams = (abi.BH_Camera * 1)(tracer.camera(0.9, 60.0, 100.0, 8, 8))
    assert lib().bh_trace_image_plane(cams, 1, None, None, 0, None, C.c_void_p(1), None) == -1
    # disk enabled but no LUT
    assert lib().bh_trace_image_plane(cams, 1, C.byref(o), None, 0, None, C.c_void_p(1), None) == -1
    # volume enabled but no volume
    o.volume_enabled = 1
    assert lib().bh_trace_image_plane(cams, 1, C.byref(o), C.c_void_p(1), 256, None,
                                      C.c_void_p(1), None) == -1


def test_mino_entries_validation_without_gpu():
    """bh_trace_disk_mino and bh_trace_disk_mino_refine reject bad arguments
    on the host, naming themselves."""
    import ctypes as C
    from blackhole_amd import abi, tracer
    from blackhole_amd._lib import lib
    L = lib()
    cams = (abi.BH_Camera * 1)(tracer.camera(0.5, 60.0, 100.0, 8, 8, 15.0))
    o = tracer.default_options(disk_enabled=1, volume_enabled=0, disk_r_out=20.0)
    for name in ("bh_trace_disk_mino", "bh_trace_disk_mino_refine"):
        fn = getattr(L, name)
        p1 = C.c_void_p(1)
        assert fn(cams, 1, C.byref(o), None, 256, p1, p1, p1, 16, None) == -1      # no LUTs
        assert name.encode() in L.bh_last_error()
        assert fn(cams, 1, C.byref(o), p1, 256, p1, p1, p1, 0, None) == -1         # no list capacity
        assert b"fallback_cap" in L.bh_last_error()
        ov = tracer.default_options(disk_enabled=1, volume_enabled=1, disk_r_out=20.0)
        assert fn(cams, 1, C.byref(ov), p1, 256, p1, p1, p1, 16, None) == -1       # volumes not supported
        assert b"thin disk only" in L.bh_last_error()
        ox = tracer.default_options(disk_enabled=1, volume_enabled=0, disk_r_out=20.0)
        assert ox.mino_exit == 0
        ox.mino_exit = 3
        assert fn(cams, 1, C.byref(ox), p1, 256, p1, p1, p1, 16, None) == -1       # unknown escape-direction mode
        assert b"mino_exit" in L.bh_last_error()


def test_measure_peaks_validation_without_gpu():
    """bh_measure_peaks rejects a NULL output before touching the device."""
    from blackhole_amd._lib import lib
    L = lib()
    assert L.bh_measure_peaks(None, None) == -1
    assert b"bh_measure_peaks" in L.bh_last_error()


def test_polarised_entry_validation_without_gpu():
    """bh_trace_image_plane_pol rejects its bad arguments on the host."""
    import ctypes as C
    from blackhole_amd import abi, tracer
    from blackhole_amd._lib import lib
    cams = (abi.BH_Camera * 1)(tracer.camera(0.5, 60.0, 100.0, 8, 8))
    o = tracer.default_options(volume_enabled=1)
    vol = abi.BH_Volume()
    vol.cells = 1
    pp = abi.BH_PolParams(0.3, 0.7, 1e-3, 0.0)
    L = lib()
    # NULL parameters, NULL Stokes buffer
    assert L.bh_trace_image_plane_pol(cams, 1, C.byref(o), None, 0, C.byref(vol), None, C.c_void_p(1),
                                      C.c_void_p(1), None) == -1
    assert L.bh_trace_image_plane_pol(cams, 1, C.byref(o), None, 0, C.byref(vol), C.byref(pp), C.c_void_p(1),
                                      None, None) == -1
    # fraction outside [0, 1]
    bad = abi.BH_PolParams(0.3, 1.5, 1e-3, 0.0)
    assert L.bh_trace_image_plane_pol(cams, 1, C.byref(o), None, 0, C.byref(vol), C.byref(bad), C.c_void_p(1),
                                      C.c_void_p(1), None) == -1
    assert b"lin_pol_frac" in L.bh_last_error()
    # no volume
    o.volume_enabled = 0
    assert L.bh_trace_image_plane_pol(cams, 1, C.byref(o), None, 0, None, C.byref(pp), C.c_void_p(1),
                                      C.c_void_p(1), None) == -1
    assert b"volume" in L.bh_last_error()


def test_struct_layouts_match_the_c_headers(tmp_path):
    """sizeof / offsetof of every ABI struct from a C compile of include/*.h
    against the ctypes mirrors."""
    import ctypes as C
    import subprocess
    from blackhole_amd import abi
    from blackhole_amd import surrogate as S
    from blackhole_amd import grmhd as G
    from blackhole_amd import luts as LM
    structs = {
        "BH_Camera": abi.BH_Camera, "BH_TraceOptions": abi.BH_TraceOptions, "BH_Volume": abi.BH_Volume,
        "BH_VolumeSet": abi.BH_VolumeSet, "BH_TuneResult": abi.BH_TuneResult, "BH_RaySpec": abi.BH_RaySpec,
        "BH_SurrogateRenderParams": S.RenderParams, "BH_SurrogateConfig": S.SurrogateConfig,
        "BH_TorusParams": abi.BH_TorusParams, "BH_GrmhdPacked": G.BH_GrmhdPacked,
        "BH_PolParams": abi.BH_PolParams, "BH_LutImage": LM.BH_LutImage, "BH_LutTableDev": LM.BH_LutTableDev,
        "BH_GrmhdStreamStats": G.BH_GrmhdStreamStats,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bhk_trace.h"', '#include "bhk_surrogate.h"',
             '#include "bh_grmhd.h"', '#include "bh_lut_manager.h"',
             "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f in cls._fields_:
            lines.append(f'  printf("{name} {f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in got:
        if not line:
            continue
        name, field, val = line.split()
        cls = structs[name]
        want = C.sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert int(val) == want, (name, field, int(val), want)


def test_frame_and_volume_set_validation_without_gpu():
    import ctypes as C
    from blackhole_amd import abi, tracer
    from blackhole_amd._lib import lib
    cams = (abi.BH_Camera * 1)(tracer.camera(1.2, 60.0, 100.0, 8, 8))  # spin out of range
    opts = tracer.default_options(disk_enabled=0)
    assert not lib().bh_frame_create(cams, 1, C.byref(opts), None, 0, None, None, C.c_void_p(1))
    assert b"spin" in lib().bh_last_error()
    cams = (abi.BH_Camera * 2)(tracer.camera(0.5, 60.0, 100.0, 8, 8), tracer.camera(0.9, 30.0, 100.0, 8, 8))
    tracer.pack_cameras(cams)
    opts = tracer.default_options(volume_enabled=1, disk_enabled=0)
    v0 = abi.BH_Volume(C.c_void_p(16), 8, 4, 2, 0, 1.1, 50.0, 5.0, 40.0, 0.7, 4.0)
    v1 = abi.BH_Volume(C.c_void_p(32), 8, 4, 4, 0, 1.1, 50.0, 5.0, 40.0, 0.7, 4.0)  # n_ph differs
    vols = (abi.BH_Volume * 2)(v0, v1)
    cv = (C.c_int32 * 2)(0, 1)
    vs = abi.BH_VolumeSet(vols, 2, 0, cv)
    assert lib().bh_trace_image_plane_multi(cams, 2, C.byref(opts), None, 0, C.byref(vs), C.c_void_p(1), None) == -1
    assert b"share grid geometry" in lib().bh_last_error()
    cv2 = (C.c_int32 * 2)(0, 5)
    vols2 = (abi.BH_Volume * 2)(v0, v0)
    vs2 = abi.BH_VolumeSet(vols2, 2, 0, cv2)
    assert lib().bh_trace_image_plane_multi(cams, 2, C.byref(opts), None, 0, C.byref(vs2), C.c_void_p(1), None) == -1
    assert b"volume index" in lib().bh_last_error()


def test_record_flag_bits_match_the_header():
    """abi.FLAG_* are the BH_FLAG_* macros of bhk_trace.h, distinct single bits."""
    txt = open(os.path.join(ROOT, "include", "bhk_trace.h")).read()
    macros = {m.group(1): int(m.group(2), 16) for m in re.finditer(r"#define BH_FLAG_(\w+)\s+0x([0-9a-fA-F]+)u", txt)}
    assert len(macros) >= 6
    for name, v in macros.items():
        assert getattr(abi, "FLAG_" + name) == v, name
        assert v & (v - 1) == 0, name
    assert len(set(macros.values())) == len(macros)
