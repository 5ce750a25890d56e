"""LUT slot registry on the GPU: tables registered into slots come back
texel for texel, re-registering replaces and frees, the GRMHD slots hold
volumes the render launch samples exactly as a volume made directly, and the
upload hands the eight handles and the background to the render launch
(bh_lut_texture_handle)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_register_read_replace_upload_and_shutdown():
    from blackhole_amd import luts
    from blackhole_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(7)
    m = luts.LutManager()
    emis = rng.uniform(0, 1, (1, 256)).astype(np.float32)
    spec = rng.uniform(0, 1, (16, 32, 4)).astype(np.float32)
    galaxy = rng.uniform(0, 1, (6, 1, 8, 8, 4)).astype(np.float32)
    vol = rng.uniform(0, 1, (8, 6, 10, 4)).astype(np.float32)  # phi, theta, r, RGBA
    vol_r = rng.uniform(0, 1, (8, 6, 10, 4)).astype(np.float32)
    m.register(luts.EMISSIVITY, emis)
    m.register(luts.SPECTRAL, spec)
    m.register(luts.GALAXY, galaxy)
    m.register(luts.GRMHD, vol)
    m.register(luts.GRMHD_RIGHT, vol_r)
    m.register_background(rng.uniform(0, 1, (4, 8, 4)).astype(np.float32))
    print("[luts] registered 5 slots and the background", flush=True)
    handles = [m.get_tex(s) for s in range(8)]
    live = [h for h in handles if h]
    assert len(live) == 5 and len(set(live)) == 5 and m.background_tex() not in live
    for slot, arr, dims in ((luts.EMISSIVITY, emis, (256, 1, 1, 1, 1)), (luts.SPECTRAL, spec, (32, 16, 1, 4, 1)),
                            (luts.GALAXY, galaxy, (8, 8, 1, 4, 6)), (luts.GRMHD, vol, (10, 6, 8, 4, 1))):
        meta, data = luts.read(m.get_tex(slot))
        assert (meta.width, meta.height, meta.depth, meta.channels, meta.layers) == dims, slot
        assert np.array_equal(data, arr.reshape(-1)), slot
    # the GRMHD slots: the render launch's sample equals that of volumes made directly
    from blackhole_amd import grmhd as G
    pts = rng.uniform(-40, 40, (200, 3)).astype(np.float32)
    t0, t1 = G.Volume(vol), G.Volume(vol_r)
    try:
        want = G.sample(t0.handle, t1.handle, 0.3, 1.0, 50.0, pts)
        got = G.sample(m.get_tex(luts.GRMHD), m.get_tex(luts.GRMHD_RIGHT), 0.3, 1.0, 50.0, pts)
        assert np.array_equal(got, want) and np.abs(want).max() > 0
    finally:
        t0.close()
        t1.close()
    # re-registering replaces the table (the old handle is freed)
    old = m.get_tex(luts.EMISSIVITY)
    emis2 = emis[:, ::-1].copy()
    m.register(luts.EMISSIVITY, emis2)
    new = m.get_tex(luts.EMISSIVITY)
    meta, data = luts.read(new)
    assert np.array_equal(data, emis2.reshape(-1))
    if new != old:
        assert L.bh_lut_read(old, C.byref(luts.BH_LutImage()), None, 0) == -1
    # upload: the render launch's slots hold the manager's handles
    m.upload()
    for s in range(8):
        assert L.bh_lut_texture_handle(s) == m.get_tex(s), s
    assert L.bh_lut_texture_handle(8) == m.background_tex()
    m.unregister(luts.SPECTRAL)
    assert m.get_tex(luts.SPECTRAL) == 0
    m.shutdown()
    assert all(m.get_tex(s) == 0 for s in range(8)) and m.background_tex() == 0
    m.upload()
    assert all(L.bh_lut_texture_handle(s) == 0 for s in range(9))
    m.close()
