"""LUT slot registry (bh_lut_manager.h), host side: slot and geometry
validation of lutRegister (src/cuda/lut_manager.h:67-99) and the empty
manager's handles, without a GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from blackhole_amd import luts
from blackhole_amd._lib import lib


def test_slot_and_geometry_validation_and_empty_handles():
    L = luts._sigs(lib())
    m = L.bh_lut_manager_create()
    assert m
    try:
        tab, keep = luts.image(np.zeros((1, 256), np.float32))
        for slot in (-1, 8, 99):
            assert L.bh_lut_register(m, slot, C.byref(tab)) == -1
            assert f"invalid slot {slot}".encode() in L.bh_last_error()
            assert L.bh_lut_get_tex(m, slot) == 0
        vol, keep2 = luts.image(np.zeros((4, 4, 8, 4), np.float32))
        # a 2-D slot refuses a volume, a GRMHD slot refuses a 2-D single-channel table,
        # the galaxy slot takes a 6-layer cubemap only
        assert L.bh_lut_register(m, luts.EMISSIVITY, C.byref(vol)) == -1 and b"2-D" in L.bh_last_error()
        assert L.bh_lut_register(m, luts.GRMHD, C.byref(tab)) == -1 and b"RGBA32F" in L.bh_last_error()
        assert L.bh_lut_register(m, luts.GALAXY, C.byref(tab)) == -1 and b"cubemap" in L.bh_last_error()
        bad = luts.BH_LutImage(tab.host, 16, 16, 1, 3, 1, 0)
        assert L.bh_lut_register(m, luts.SPECTRAL, C.byref(bad)) == -1 and b"channels" in L.bh_last_error()
        for s in range(8):
            assert L.bh_lut_get_tex(m, s) == 0
            L.bh_lut_unregister(m, s)  # no effect on an empty slot
        assert L.bh_lut_get_background_tex(m) == 0
        assert L.bh_lut_register_background(m, C.byref(vol)) == -1
        L.bh_lut_shutdown(m)
        assert L.bh_lut_get_tex(None, 0) == 0
        del keep, keep2
    finally:
        L.bh_lut_manager_destroy(m)
    L.bh_lut_manager_destroy(None)
