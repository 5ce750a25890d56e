"""LUT slot registry of the render path (Python over bh_lut_manager.h).

Mirrors the reference's LutManager (src/cuda/lut_manager.h:15-134): eight
slots (BhLutSlot) plus a standalone background, register / unregister /
get / shutdown.  Tables are registered from host arrays and held on the
device; the GRMHD slots hold volumes the render launch samples.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

EMISSIVITY, REDSHIFT, SPECTRAL, GRB, GALAXY, GRMHD, SYNCH_G, GRMHD_RIGHT = range(8)
SLOT_COUNT = 8


class BH_LutImage(C.Structure):
    _fields_ = [("host", C.POINTER(C.c_float)), ("width", C.c_int32), ("height", C.c_int32),
                ("depth", C.c_int32), ("channels", C.c_int32), ("layers", C.c_int32), ("pad0", C.c_int32)]


class BH_LutTableDev(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("depth", C.c_int32),
                ("channels", C.c_int32), ("layers", C.c_int32), ("slot", C.c_int32)]


def _sigs(L):
    if getattr(L, "_lut_sigs", False):
        return L
    L.bh_lut_manager_create.restype = C.c_void_p
    L.bh_lut_manager_create.argtypes = []
    L.bh_lut_register.argtypes = [C.c_void_p, C.c_int, C.POINTER(BH_LutImage)]
    L.bh_lut_register.restype = C.c_int
    L.bh_lut_unregister.argtypes = [C.c_void_p, C.c_int]
    L.bh_lut_unregister.restype = None
    L.bh_lut_get_tex.argtypes = [C.c_void_p, C.c_int]
    L.bh_lut_get_tex.restype = C.c_ulonglong
    L.bh_lut_register_background.argtypes = [C.c_void_p, C.POINTER(BH_LutImage)]
    L.bh_lut_register_background.restype = C.c_int
    L.bh_lut_unregister_background.argtypes = [C.c_void_p]
    L.bh_lut_unregister_background.restype = None
    L.bh_lut_get_background_tex.argtypes = [C.c_void_p]
    L.bh_lut_get_background_tex.restype = C.c_ulonglong
    L.bh_lut_shutdown.argtypes = [C.c_void_p]
    L.bh_lut_shutdown.restype = None
    L.bh_lut_manager_destroy.argtypes = [C.c_void_p]
    L.bh_lut_manager_destroy.restype = None
    L.bh_lut_upload.argtypes = [C.c_void_p]
    L.bh_lut_upload.restype = C.c_int
    L.bh_lut_read.argtypes = [C.c_ulonglong, C.POINTER(BH_LutImage), C.POINTER(C.c_float), C.c_int64]
    L.bh_lut_read.restype = C.c_int
    L._lut_sigs = True
    return L


def image(arr: np.ndarray, layers: int = 1) -> tuple[BH_LutImage, np.ndarray]:
    """BH_LutImage over a float32 array shaped [layers][depth][height][width][channels]
    (leading axes may be omitted); returns the struct and the array it points to."""
    a = np.ascontiguousarray(arr, dtype=np.float32)
    shp = a.shape
    ch = shp[-1] if a.ndim >= 3 and shp[-1] in (1, 4) else 1
    dims = shp[:-1] if a.ndim >= 3 and shp[-1] in (1, 4) else shp
    dims = (1,) * (4 - len(dims)) + tuple(dims) if len(dims) < 4 else tuple(dims)
    lay, d, h, w = dims
    img = BH_LutImage(a.ctypes.data_as(C.POINTER(C.c_float)), w, h, d, ch, lay if lay > 1 else layers, 0)
    return img, a


class LutManager:
    """Owns a BH_LutManager; tables stay on the device until unregistered."""

    def __init__(self):
        self.L = _sigs(lib())
        self.h = self.L.bh_lut_manager_create()
        if not self.h:
            check(-1, "bh_lut_manager_create")

    def register(self, slot: int, arr: np.ndarray, layers: int = 1) -> None:
        img, keep = image(arr, layers)
        check(self.L.bh_lut_register(self.h, slot, C.byref(img)), "bh_lut_register")
        del keep

    def unregister(self, slot: int) -> None:
        self.L.bh_lut_unregister(self.h, slot)

    def get_tex(self, slot: int) -> int:
        return int(self.L.bh_lut_get_tex(self.h, slot))

    def register_background(self, arr: np.ndarray) -> None:
        img, keep = image(arr)
        check(self.L.bh_lut_register_background(self.h, C.byref(img)), "bh_lut_register_background")
        del keep

    def background_tex(self) -> int:
        return int(self.L.bh_lut_get_background_tex(self.h))

    def upload(self) -> None:
        check(self.L.bh_lut_upload(self.h), "bh_lut_upload")

    def shutdown(self) -> None:
        self.L.bh_lut_shutdown(self.h)

    def close(self) -> None:
        if self.h:
            self.L.bh_lut_manager_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read(handle: int) -> tuple[BH_LutImage, np.ndarray]:
    """Geometry and texels of a registered table (device -> host)."""
    L = _sigs(lib())
    meta = BH_LutImage()
    check(L.bh_lut_read(handle, C.byref(meta), None, 0), "bh_lut_read")
    n = meta.width * meta.height * meta.depth * meta.channels * meta.layers
    out = np.empty(n, dtype=np.float32)
    check(L.bh_lut_read(handle, C.byref(meta), out.ctypes.data_as(C.POINTER(C.c_float)), n), "bh_lut_read")
    return meta, out
