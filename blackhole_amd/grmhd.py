"""GRMHD packed volumes (nubhlight_pack format) over include/bh_grmhd.h.

load_packed / timeseries_info / load_frame / frame_pair wrap the C loader;
write_packed and write_timeseries produce the same format (metadata JSON +
RGBA32F binary, x = r fastest) from numpy arrays, for synthetic datasets;
Volume owns a device volume usable in the render launch's grmhd slots.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from ._lib import check, lib


class BH_GrmhdPacked(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("depth", C.c_int32),
                ("checksum_verified", C.c_int32), ("r_min", C.c_float), ("r_max", C.c_float),
                ("min", C.c_float * 4), ("max", C.c_float * 4), ("channels", (C.c_char * 32) * 4),
                ("checksum", C.c_uint64), ("floats", C.c_int64)]


def _sigs():
    L = lib()
    if getattr(L, "_grmhd_sigs", False):
        return L
    vp = C.c_void_p
    L.bh_grmhd_load_packed.argtypes = [C.c_char_p, C.POINTER(BH_GrmhdPacked), vp, C.c_int64]
    L.bh_fnv1a64.argtypes = [vp, C.c_int64]
    L.bh_fnv1a64.restype = C.c_uint64
    L.bh_grmhd_timeseries_info.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int32), vp, vp, C.c_int32]
    L.bh_grmhd_load_frame.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, vp, C.c_int64]
    L.bh_grmhd_frame_pair.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
    L.bh_grmhd_texture_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_ulonglong)]
    L.bh_grmhd_texture_destroy.argtypes = [C.c_ulonglong]
    L.bh_grmhd_sample.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_float, C.c_float, C.c_float, vp, C.c_int32, vp]
    for n in ("bh_grmhd_load_packed", "bh_grmhd_timeseries_info", "bh_grmhd_load_frame", "bh_grmhd_frame_pair",
              "bh_grmhd_texture_create", "bh_grmhd_texture_destroy", "bh_grmhd_sample"):
        getattr(L, n).restype = C.c_int
    L._grmhd_sigs = True
    return L


def fnv1a64(data: bytes) -> int:
    buf = C.create_string_buffer(data, len(data))
    return int(_sigs().bh_fnv1a64(C.cast(buf, C.c_void_p), len(data)))


def load_packed(json_path: str, data: bool = True):
    """(metadata dict, float32 [depth, height, width, 4] or None)."""
    L = _sigs()
    m = BH_GrmhdPacked()
    check(L.bh_grmhd_load_packed(json_path.encode(), C.byref(m), None, 0), "bh_grmhd_load_packed")
    arr = None
    if data:
        arr = np.empty((m.depth, m.height, m.width, 4), dtype=np.float32)
        check(L.bh_grmhd_load_packed(json_path.encode(), C.byref(m), arr.ctypes.data_as(C.c_void_p), arr.size),
              "bh_grmhd_load_packed")
    meta = {"dims": (m.width, m.height, m.depth), "r_min": m.r_min, "r_max": m.r_max,
            "min": list(m.min), "max": list(m.max), "channels": [bytes(c.value).decode() for c in m.channels],
            "checksum": int(m.checksum), "checksum_verified": bool(m.checksum_verified)}
    return meta, arr


def write_packed(json_path: str, rgba: np.ndarray, channels=("rho", "u", "v1", "v2"), r_min=None, r_max=None,
                 bin_name: str | None = None, checksum: bool = True) -> None:
    """rgba: float32 [depth (phi), height (theta), width (r), 4]."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    d, h, w, c = rgba.shape
    assert c == 4
    bin_name = bin_name or os.path.splitext(os.path.basename(json_path))[0] + ".bin"
    raw = rgba.tobytes()
    with open(os.path.join(os.path.dirname(os.path.abspath(json_path)), bin_name), "wb") as f:
        f.write(raw)
    flat = rgba.reshape(-1, 4)
    meta = {"schema_version": 1, "source": "synthetic", "layout": "channels-last", "format": "RGBA32F",
            "bin": bin_name, "grid_dims": [w, h, d], "channels": list(channels),
            "min": [float(x) for x in flat.min(0)], "max": [float(x) for x in flat.max(0)]}
    if checksum:
        meta["checksum_fnv1a64"] = "%016x" % fnv1a64(raw)
    if r_min is not None:
        meta["r_min"] = float(r_min)
    if r_max is not None:
        meta["r_max"] = float(r_max)
    with open(json_path, "w") as f:
        json.dump(meta, f, indent=2)


def write_timeseries(json_path: str, bin_path: str, frames, times, channels=("rho", "u", "v1", "v2")) -> None:
    """frames: list of float32 [depth, height, width, 4] arrays of equal shape."""
    d, h, w, _ = frames[0].shape
    size = d * h * w * 16
    with open(bin_path, "wb") as f:
        for fr in frames:
            f.write(np.ascontiguousarray(fr, dtype=np.float32).tobytes())
    meta = {"version": "1.0-multi", "frameCount": len(frames), "frameSize": size,
            "frames": [{"file": "frame%04d.h5" % i, "time": float(t), "offset": i * size} for i, t in enumerate(times)],
            "frameMetadata": {"gridDims": [w, h, d], "channels": list(channels), "format": "RGBA32F"}}
    with open(json_path, "w") as f:
        json.dump(meta, f, indent=2)


def timeseries_info(json_path: str):
    L = _sigs()
    n, fb = C.c_int32(), C.c_int64()
    dims = (C.c_int32 * 3)()
    times = np.zeros(65536)
    offs = np.zeros(65536, dtype=np.int64)
    check(L.bh_grmhd_timeseries_info(json_path.encode(), C.byref(n), C.byref(fb), dims,
                                     times.ctypes.data_as(C.c_void_p), offs.ctypes.data_as(C.c_void_p), 65536),
          "bh_grmhd_timeseries_info")
    return {"frames": n.value, "frame_bytes": fb.value, "dims": tuple(dims), "times": times[:n.value].copy(),
            "offsets": offs[:n.value].copy()}


def load_frame(json_path: str, bin_path: str, frame: int) -> np.ndarray:
    info = timeseries_info(json_path)
    w, h, d = info["dims"]
    out = np.empty((d, h, w, 4), dtype=np.float32)
    check(_sigs().bh_grmhd_load_frame(json_path.encode(), bin_path.encode(), frame, out.ctypes.data_as(C.c_void_p),
                                      out.size), "bh_grmhd_load_frame")
    return out


def frame_pair(times, t: float):
    times = np.ascontiguousarray(times, dtype=np.float64)
    i0, a = C.c_int32(), C.c_float()
    check(_sigs().bh_grmhd_frame_pair(times.ctypes.data_as(C.c_void_p), len(times), t, C.byref(i0), C.byref(a)),
          "bh_grmhd_frame_pair")
    return i0.value, a.value


class Volume:
    """A device volume (bh_grmhd_texture_create) for the render launch's grmhd slots."""

    def __init__(self, rgba: np.ndarray):
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        d, h, w, _ = rgba.shape
        self.handle = C.c_ulonglong()
        check(_sigs().bh_grmhd_texture_create(rgba.ctypes.data_as(C.c_void_p), w, h, d, C.byref(self.handle)),
              "bh_grmhd_texture_create")
        self.handle = self.handle.value

    def close(self) -> None:
        if self.handle:
            check(_sigs().bh_grmhd_texture_destroy(self.handle), "bh_grmhd_texture_destroy")
            self.handle = 0


def sample(left: int, right: int, alpha: float, r_min: float, r_max: float, xyz: np.ndarray) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    out = np.zeros((xyz.shape[0], 4), dtype=np.float32)
    check(_sigs().bh_grmhd_sample(left, right, alpha, r_min, r_max, xyz.ctypes.data_as(C.c_void_p), xyz.shape[0],
                                  out.ctypes.data_as(C.c_void_p)), "bh_grmhd_sample")
    return out


class BH_GrmhdStreamStats(C.Structure):
    _fields_ = [("hits", C.c_uint64), ("misses", C.c_uint64), ("resident_bytes", C.c_int64),
                ("budget_bytes", C.c_int64), ("resident_frames", C.c_int32), ("queue_depth", C.c_int32),
                ("current_frame", C.c_int32), ("frame_count", C.c_int32)]


def _stream_sigs():
    L = _sigs()
    if getattr(L, "_grmhd_stream_sigs", False):
        return L
    L.bh_grmhd_streamer_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, C.c_int32]
    L.bh_grmhd_streamer_create.restype = C.c_void_p
    L.bh_grmhd_streamer_destroy.argtypes = [C.c_void_p]
    L.bh_grmhd_streamer_destroy.restype = None
    L.bh_grmhd_streamer_seek.argtypes = [C.c_void_p, C.c_int32]
    L.bh_grmhd_streamer_get_frame.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_ulonglong)]
    L.bh_grmhd_streamer_frame_pair.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.POINTER(C.c_ulonglong),
                                               C.POINTER(C.c_ulonglong), C.POINTER(C.c_float)]
    L.bh_grmhd_streamer_stats.argtypes = [C.c_void_p, C.POINTER(BH_GrmhdStreamStats)]
    for n in ("bh_grmhd_streamer_seek", "bh_grmhd_streamer_get_frame", "bh_grmhd_streamer_frame_pair",
              "bh_grmhd_streamer_stats"):
        getattr(L, n).restype = C.c_int
    L._grmhd_stream_sigs = True
    return L


class Streamer:
    """GRMHD time series streamed into device volumes (bh_grmhd_streamer_*,
    the reference's GRMHDStreamer): an LRU frame cache in HBM under a byte
    budget, a loader thread, and prefetch of the next frames on a seek."""

    def __init__(self, json_path: str, bin_path: str, cache_bytes: int, prefetch: int = 3):
        self.L = _stream_sigs()
        self.h = self.L.bh_grmhd_streamer_create(json_path.encode(), bin_path.encode(), int(cache_bytes), prefetch)
        if not self.h:
            check(-1, "bh_grmhd_streamer_create")

    def seek(self, frame: int) -> None:
        check(self.L.bh_grmhd_streamer_seek(self.h, frame), "bh_grmhd_streamer_seek")

    def get_frame(self, offset: int = 0, wait: bool = False) -> int:
        """Device volume of the current (offset 0) or next (1) frame, 0 when not resident."""
        t = C.c_ulonglong()
        rc = self.L.bh_grmhd_streamer_get_frame(self.h, offset, 1 if wait else 0, C.byref(t))
        if rc < 0:
            check(rc, "bh_grmhd_streamer_get_frame")
        return int(t.value)

    def frame_pair(self, t: float, wait: bool = True):
        """(left volume, right volume, alpha) around simulation time t."""
        a, b, al = C.c_ulonglong(), C.c_ulonglong(), C.c_float()
        rc = self.L.bh_grmhd_streamer_frame_pair(self.h, t, 1 if wait else 0, C.byref(a), C.byref(b), C.byref(al))
        if rc < 0:
            check(rc, "bh_grmhd_streamer_frame_pair")
        return int(a.value), int(b.value), float(al.value)

    def stats(self) -> dict:
        s = BH_GrmhdStreamStats()
        check(self.L.bh_grmhd_streamer_stats(self.h, C.byref(s)), "bh_grmhd_streamer_stats")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self) -> None:
        if self.h:
            self.L.bh_grmhd_streamer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
