"""GRMHD packed-volume ingestion (include/bh_grmhd.h) on the CPU: the FNV-1a64
checksum against its published test vectors, a round trip through the
nubhlight_pack format (the fixture of the reference's pack test:
value = i0 + 0.1 i1 + 0.01 i2 + 0.5 c on a 2 x 2 x 2 grid), every validation
error of the reference loader, the legacy metadata forms, the time-series
metadata, and the frame-pair rule of temporal interpolation."""
from __future__ import annotations

import json

import numpy as np
import pytest

from blackhole_amd import grmhd as G
from blackhole_amd._lib import lib


def test_fnv1a64_published_vectors():
    # Fowler-Noll-Vo test vectors (64-bit FNV-1a)
    assert G.fnv1a64(b"") == 0xCBF29CE484222325
    assert G.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert G.fnv1a64(b"foobar") == 0x85944171F73967E8


def _fixture():
    d = np.zeros((2, 2, 2, 4), dtype=np.float32)  # [i0, i1, i2, c] as the packer's grid order
    for i0 in range(2):
        for i1 in range(2):
            for i2 in range(2):
                for c in range(4):
                    d[i0, i1, i2, c] = np.float32(i0 + i1 * 0.1 + i2 * 0.01 + c * 0.5)
    return d


def test_round_trip_of_the_reference_fixture(tmp_path):
    data = _fixture()
    path = str(tmp_path / "grmhd_pack.json")
    G.write_packed(path, data, channels=("RHO", "UU", "U1", "U2"))
    meta, arr = G.load_packed(path)
    assert meta["dims"] == (2, 2, 2) and meta["channels"] == ["RHO", "UU", "U1", "U2"]
    assert meta["checksum_verified"] and meta["r_min"] == 1.0 and meta["r_max"] == 50.0
    assert np.array_equal(arr, data)
    flat = data.reshape(-1, 4)
    assert np.allclose(meta["min"], flat.min(0)) and np.allclose(meta["max"], flat.max(0))
    assert meta["checksum"] == G.fnv1a64(data.tobytes())


def _expect_error(path, text):
    with pytest.raises(RuntimeError) as e:
        G.load_packed(path)
    assert text in str(e.value), str(e.value)


def test_validation_errors(tmp_path):
    data = _fixture()
    path = str(tmp_path / "g.json")
    G.write_packed(path, data)
    meta = json.load(open(path))
    # checksum mismatch
    raw = bytearray(data.tobytes())
    raw[5] ^= 1
    (tmp_path / "g.bin").write_bytes(bytes(raw))
    _expect_error(path, "checksum mismatch")
    # size mismatch
    (tmp_path / "g.bin").write_bytes(data.tobytes()[:-16])
    _expect_error(path, "data size mismatch")
    (tmp_path / "g.bin").write_bytes(data.tobytes())
    for key, value, text in (("format", "R32F", "unsupported texture format"),
                             ("grid_dims", [2, 2], "grid_dims"),
                             ("channels", ["a", "b", "c"], "expected 4 channels")):
        m = dict(meta)
        m[key] = value
        json.dump(m, open(path, "w"))
        _expect_error(path, text)
    m = dict(meta)
    del m["bin"]
    json.dump(m, open(path, "w"))
    _expect_error(path, "missing 'bin'")
    (tmp_path / "bad.json").write_text("{ not json")
    _expect_error(str(tmp_path / "bad.json"), "invalid JSON")
    _expect_error(str(tmp_path / "absent.json"), "cannot read metadata")


def test_legacy_metadata_forms(tmp_path):
    data = _fixture()
    path = str(tmp_path / "g.json")
    G.write_packed(path, data, checksum=False, r_min=2.0, r_max=30.0)
    m = json.load(open(path))
    m["binary_file"] = m.pop("bin")
    m["channels"] = [{"name": n, "min": float(i), "max": float(10 + i)} for i, n in enumerate("abcd")]
    del m["min"], m["max"]
    json.dump(m, open(path, "w"))
    meta, arr = G.load_packed(path)
    assert not meta["checksum_verified"] and np.array_equal(arr, data)
    assert meta["channels"] == ["a", "b", "c", "d"] and meta["min"] == [0, 1, 2, 3] and meta["max"] == [10, 11, 12, 13]
    assert meta["r_min"] == 2.0 and meta["r_max"] == 30.0


def test_timeseries_and_frame_pair(tmp_path):
    rng = np.random.default_rng(3)
    frames = [rng.random((3, 4, 5, 4), dtype=np.float32) for _ in range(3)]
    js, bn = str(tmp_path / "ts.json"), str(tmp_path / "ts.bin")
    G.write_timeseries(js, bn, frames, [0.0, 1.0, 2.5])
    info = G.timeseries_info(js)
    assert info["frames"] == 3 and info["dims"] == (5, 4, 3) and info["frame_bytes"] == 3 * 4 * 5 * 16
    assert np.array_equal(info["times"], [0.0, 1.0, 2.5])
    for k in range(3):
        assert np.array_equal(G.load_frame(js, bn, k), frames[k])
    times = [0.0, 1.0, 2.5]
    assert G.frame_pair(times, 0.5) == (0, 0.5)
    i0, a = G.frame_pair(times, 2.0)
    assert i0 == 1 and abs(a - 1.0 / 1.5) < 1e-7
    assert G.frame_pair(times, -1.0) == (0, 0.0) and G.frame_pair(times, 9.0) == (1, 1.0)
    assert G.frame_pair([4.0], 7.0) == (0, 0.0)
    with pytest.raises(RuntimeError):
        G.load_frame(js, bn, 3)


def test_volume_calls_validate_without_gpu():
    import ctypes as C
    L = G._sigs()
    h = C.c_ulonglong()
    assert L.bh_grmhd_texture_create(None, 2, 2, 2, C.byref(h)) == -1
    assert L.bh_grmhd_texture_destroy(12345) == -1
    assert b"not a volume" in lib().bh_last_error()
    xyz = np.zeros(3, dtype=np.float32)
    out = np.zeros(4, dtype=np.float32)
    assert L.bh_grmhd_sample(777, 0, 0.0, 1.0, 50.0, xyz.ctypes.data_as(C.c_void_p), 1,
                             out.ctypes.data_as(C.c_void_p)) == -1
