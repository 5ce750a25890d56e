"""Host-side interface of the image-plane tracer (Python over the C ABI).

Mirrors the reference's image-plane entry (bhbCudaTraceGeodesics,
src/blender_bridge/blender_bridge.h:348-350: spin, observer radius,
inclination, alpha/beta window and pixel counts in; per-pixel results out)
on top of ``bh_trace_image_plane`` (include/bhk_trace.h).  Device memory and
streams come from PyTorch (plumbing only); every computation runs in the HIP
library, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import abi
from ._lib import check, lib

def default_options(**overrides) -> abi.BH_TraceOptions:
    o = abi.BH_TraceOptions()
    check(lib().bh_trace_options_default(C.byref(o)), "bh_trace_options_default")
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"BH_TraceOptions has no field {k!r}")
        setattr(o, k, v)
    return o


def camera(spin: float, theta_obs_deg: float, r_obs: float, width: int, height: int,
           half_extent: float = 15.0, record_offset: int = 0) -> abi.BH_Camera:
    return abi.BH_Camera(spin, math.radians(theta_obs_deg), r_obs, -half_extent, half_extent,
                         -half_extent, half_extent, width, height, record_offset)


def pack_cameras(cams: Sequence[abi.BH_Camera]) -> int:
    """Assign consecutive record ranges; returns the total record count."""
    off = 0
    for c in cams:
        c.record_offset = off
        off += c.width * c.height
    return off


def isco(spin: float) -> float:
    return lib().bh_isco_prograde(spin)


def outer_horizon(spin: float) -> float:
    return lib().bh_outer_horizon(spin)


def disk_luts(cams: Sequence[abi.BH_Camera], r_out: float, n: int = 256) -> np.ndarray:
    """Per-camera normalised NT flux tables over [r_isco(spin), r_out] (host)."""
    out = np.zeros((len(cams), n))
    for i, c in enumerate(cams):
        row = np.zeros(n)
        check(lib().bh_disk_lut_build(c.spin, isco(c.spin), r_out, n,
                                      row.ctypes.data_as(C.POINTER(C.c_double))),
              "bh_disk_lut_build")
        out[i] = row
    return out


def torus_params(spin: float, n_r: int = 256, n_th: int = 128, n_ph: int = 64,
                 r_hi: float = 50.0, seed: int = 0x5EED0001, **kw) -> abi.BH_TorusParams:
    p = abi.BH_TorusParams()
    p.spin = spin
    p.r_in = 6.0
    p.r_max = 12.0
    p.r_lo = outer_horizon(spin)
    p.r_hi = r_hi
    p.n_r, p.n_th, p.n_ph = n_r, n_th, n_ph
    p.sigma = 0.5
    p.spectral_index = -11.0 / 3.0
    p.seed = seed
    p.n_e_max = 1.0
    p.theta_e_max = 30.0
    p.b_max = 30.0
    p.gamma_ad = 4.0 / 3.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def generate_torus(p: abi.BH_TorusParams):
    """Host torus grid: returns (cells float32[n_ph, n_th, n_r, 4] holding n_e, Theta_e,
    |B| and K_2(1/Theta_e)/(2 Theta_e^2), BH_Volume without the device pointer)."""
    cells = np.zeros((p.n_ph, p.n_th, p.n_r, 4), dtype=np.float32)
    vol = abi.BH_Volume()
    check(lib().bh_volume_generate_torus(C.byref(p), cells.ctypes.data_as(C.POINTER(C.c_float)),
                                         C.byref(vol)), "bh_volume_generate_torus")
    return cells, vol


@dataclass
class DeviceScene:
    """Device-resident inputs of one trace (tensors keep the memory alive)."""
    cams: list
    opts: abi.BH_TraceOptions
    n_records: int
    luts: Optional[object] = None      # torch float64 [n_cams, lut_n] on cuda
    lut_n: int = 0
    cells: Optional[object] = None     # torch float32 [.., 4] on cuda
    vol: Optional[abi.BH_Volume] = None


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: the tracer has no CPU path")
    return torch


def upload_scene(cams, opts, luts: Optional[np.ndarray] = None, cells: Optional[np.ndarray] = None,
                 vol: Optional[abi.BH_Volume] = None, device: str = "cuda:0") -> DeviceScene:
    torch = _torch()
    n = max(c.record_offset + c.width * c.height for c in cams)
    sc = DeviceScene(cams=list(cams), opts=opts, n_records=n)
    if luts is not None:
        sc.luts = torch.from_numpy(np.ascontiguousarray(luts, dtype=np.float64)).to(device)
        sc.lut_n = luts.shape[-1]
    if cells is not None:
        sc.cells = torch.from_numpy(np.ascontiguousarray(cells)).to(device)
        sc.vol = abi.BH_Volume()
        C.pointer(sc.vol)[0] = vol
        sc.vol.cells = sc.cells.data_ptr()
    return sc


def alloc_records(n: int, device: str = "cuda:0"):
    torch = _torch()
    return torch.empty((n, 128), dtype=torch.uint8, device=device)


def launch(sc: DeviceScene, out, stream=None) -> None:
    """Asynchronous trace of every pixel of sc.cams into out (uint8 [n, 128] cuda)."""
    torch = _torch()
    assert out.dtype == torch.uint8 and out.is_cuda and out.shape[1] == 128
    assert out.shape[0] >= sc.n_records
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().bh_trace_image_plane(
        cam_arr, len(sc.cams), C.byref(sc.opts),
        C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
        C.byref(sc.vol) if sc.vol is not None else None,
        C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream),
    )
    check(rc, "bh_trace_image_plane")


def launch_pol(sc: DeviceScene, out, stokes, evpa_b: float, lin_pol_frac: float,
               faraday_per_ne: float, stream=None) -> None:
    """launch() with the polarised transfer (FP64, one volume): the same
    records into out and (I, Q, U, V) per record into stokes (float64 cuda
    [n, 4])."""
    torch = _torch()
    assert out.dtype == torch.uint8 and out.is_cuda and out.shape[1] == 128
    assert out.shape[0] >= sc.n_records
    assert stokes.dtype == torch.float64 and stokes.is_cuda and stokes.is_contiguous()
    assert stokes.shape[0] >= sc.n_records and stokes.shape[1] == 4
    if sc.vol is None:
        raise ValueError("polarised transfer needs a volume scene")
    pp = abi.BH_PolParams(evpa_b, lin_pol_frac, faraday_per_ne, 0.0)
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    s = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().bh_trace_image_plane_pol(
        cam_arr, len(sc.cams), C.byref(sc.opts),
        C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
        C.byref(sc.vol), C.byref(pp), C.c_void_p(out.data_ptr()), C.c_void_p(stokes.data_ptr()),
        C.c_void_p(s.cuda_stream),
    )
    check(rc, "bh_trace_image_plane_pol")


def launch_ids(sc: DeviceScene, ids, count, out, stream=None, compact: bool = False) -> None:
    """Asynchronous trace of the pixels listed in ids (int64 cuda) with the
    list length in count (int64/uint64 cuda, one element), records to out[id]
    (compact: list entry q to out[q])."""
    torch = _torch()
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    s = stream if stream is not None else torch.cuda.current_stream()
    fn = lib().bh_trace_pixel_list_compact if compact else lib().bh_trace_pixel_list
    rc = fn(
        cam_arr, len(sc.cams), C.byref(sc.opts),
        C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
        C.byref(sc.vol) if sc.vol is not None else None,
        C.c_void_p(ids.data_ptr()), C.c_void_p(count.data_ptr()), int(ids.numel()),
        C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream))
    check(rc, "bh_trace_pixel_list_compact" if compact else "bh_trace_pixel_list")


def trace_disk_mino(sc: DeviceScene, out, fb_ids, fb_count, stream=None, opts=None) -> None:
    """Asynchronous thin-disk records of every pixel from the Mino-time
    crossing solver (bh_trace_disk_mino) into out; ids of unresolved rays to
    fb_ids (int64 cuda) with their number in fb_count (int64 cuda, one
    element, zeroed by the call).  opts: the scene's options by default."""
    torch = _torch()
    assert out.dtype == torch.uint8 and out.is_cuda and out.shape[1] == 128
    assert out.shape[0] >= sc.n_records and sc.luts is not None
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    s = stream if stream is not None else torch.cuda.current_stream()
    o = sc.opts if opts is None else opts
    rc = lib().bh_trace_disk_mino(cam_arr, len(sc.cams), C.byref(o), C.c_void_p(sc.luts.data_ptr()),
                                  sc.lut_n, C.c_void_p(out.data_ptr()), C.c_void_p(fb_ids.data_ptr()),
                                  C.c_void_p(fb_count.data_ptr()), int(fb_ids.numel()), C.c_void_p(s.cuda_stream))
    check(rc, "bh_trace_disk_mino")


def trace_rays(specs, opts, out=None, stream=None):
    """Trace explicit rays: specs float64 cuda [n, 5] = (alpha, beta, spin,
    theta_obs, r_obs); returns the uint8 [n, 128] record tensor."""
    torch = _torch()
    specs = specs.contiguous()
    assert specs.dtype == torch.float64 and specs.shape[1] == 5
    n = specs.shape[0]
    if out is None:
        out = alloc_records(n, device=specs.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().bh_trace_rays(C.c_void_p(specs.data_ptr()), n, C.byref(opts), None,
                              C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream)), "bh_trace_rays")
    return out


def records_to_numpy(out) -> np.ndarray:
    return out.cpu().numpy().view(abi.RECORD_DTYPE).reshape(-1)


def last_kernel_ms() -> float:
    return float(lib().bh_last_kernel_ms())


class VolumeSetDevice:
    """Device cells of several volumes sharing one geometry (a volume set);
    keeps the tensors and the ctypes BH_VolumeSet alive."""

    def __init__(self, cells_list, vols, cam_vol, device: str = "cuda:0"):
        torch = _torch()
        self.cells = [torch.from_numpy(np.ascontiguousarray(c)).to(device) for c in cells_list]
        self.vols = (abi.BH_Volume * len(vols))()
        for i, v in enumerate(vols):
            C.pointer(self.vols[i])[0] = v
            self.vols[i].cells = self.cells[i].data_ptr()
        self.cam_vol = (C.c_int32 * len(cam_vol))(*[int(x) for x in cam_vol])
        self.vs = abi.BH_VolumeSet(self.vols, len(vols), 0, self.cam_vol)


def launch_multi(sc: DeviceScene, vset: VolumeSetDevice, out, stream=None) -> None:
    """Asynchronous trace of every pixel of sc.cams, camera i sampling volume
    vset.cam_vol[i] (bh_trace_image_plane_multi)."""
    torch = _torch()
    cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().bh_trace_image_plane_multi(
        cam_arr, len(sc.cams), C.byref(sc.opts),
        C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
        C.byref(vset.vs), C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream)), "bh_trace_image_plane_multi")


class Frame:
    """A prepared frame (bh_frame_*): launches without host work, optionally
    replayed from a captured hipGraph, with an autotuned launch shape."""

    def __init__(self, sc: DeviceScene, out, vset: Optional[VolumeSetDevice] = None):
        self.sc, self.out, self.vset = sc, out, vset
        cam_arr = (abi.BH_Camera * len(sc.cams))(*sc.cams)
        self.h = lib().bh_frame_create(
            cam_arr, len(sc.cams), C.byref(sc.opts),
            C.c_void_p(sc.luts.data_ptr()) if sc.luts is not None else None, sc.lut_n,
            C.byref(sc.vol) if (sc.vol is not None and vset is None) else None,
            C.byref(vset.vs) if vset is not None else None, C.c_void_p(out.data_ptr()))
        if not self.h:
            check(-1, "bh_frame_create")

    def _s(self, stream):
        torch = _torch()
        return C.c_void_p((stream if stream is not None else torch.cuda.current_stream()).cuda_stream)

    def launch(self, stream=None) -> None:
        check(lib().bh_frame_launch(self.h, self._s(stream)), "bh_frame_launch")

    def graph_launch(self, stream=None) -> None:
        check(lib().bh_frame_graph_launch(self.h, self._s(stream)), "bh_frame_graph_launch")

    def set_launch_shape(self, blocks_per_cu: int, refill_batch: int) -> None:
        check(lib().bh_frame_set_launch_shape(self.h, blocks_per_cu, refill_batch), "bh_frame_set_launch_shape")

    def autotune(self, sample_stride: int = 8, stream=None) -> abi.BH_TuneResult:
        r = abi.BH_TuneResult()
        check(lib().bh_frame_autotune(self.h, sample_stride, C.byref(r), self._s(stream)), "bh_frame_autotune")
        return r

    def close(self) -> None:
        if self.h:
            lib().bh_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_params(rs: float = 2.0, spin: float = 0.9, cam_pos=(0.0, -20.0, 12.0), look_at=(0.0, 0.0, 0.0),
                  fov_scale: float = 0.35, width: int = 64, height: int = 48, max_steps: int = 20000,
                  max_dist: float = 400.0, isco: float = 0.0, **kw) -> abi.BH_LaunchParams:
    """A BH_LaunchParams for a pinhole camera at cam_pos looking at look_at
    (z = spin axis); basis columns [right, up, forward] with up toward +z."""
    p = abi.BH_LaunchParams()
    p.rs, p.spin, p.isco, p.step_size, p.fov_scale, p.max_dist = rs, spin, isco, 0.1, fov_scale, max_dist
    pos = np.array(cam_pos, dtype=np.float64)
    fwd = np.array(look_at, dtype=np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.array([0.0, 0.0, 1.0]))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    basis = np.concatenate([right, up, fwd])  # column-major: columns right, up, forward
    for i in range(3):
        p.cam_pos[i] = pos[i]
    for i in range(9):
        p.cam_basis[i] = basis[i]
    p.max_steps, p.width, p.height = max_steps, width, height
    p.adisk_enabled, p.redshift_enabled, p.kerr_enabled, p.adisk_lit = 1, 1, 1, 1.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def trace_launch_params(params: abi.BH_LaunchParams, variant: int = 2, out=None, stream=None):
    """Records (uint8 [w*h, 128] cuda) of a render launch (bh_trace_launch_params)."""
    torch = _torch()
    n = params.width * params.height
    if out is None:
        out = alloc_records(n)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().bh_trace_launch_params(C.byref(params), variant, C.c_void_p(out.data_ptr()),
                                       C.c_void_p(s.cuda_stream)), "bh_trace_launch_params")
    return out


def render_launch(params: abi.BH_LaunchParams, variant: int = -1, fb=None, stream=None):
    """Shaded float4 framebuffer (float32 [h, w, 4] cuda) of bh_launch_geodesic_kernel."""
    torch = _torch()
    if fb is None:
        fb = torch.empty((params.height, params.width, 4), dtype=torch.float32, device="cuda")
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().bh_launch_geodesic_kernel(C.c_void_p(fb.data_ptr()), C.byref(params), variant,
                                          C.c_void_p(s.cuda_stream)), "bh_launch_geodesic_kernel")
    return fb
