"""GRMHD time-series streaming into device memory (bh_grmhd_streamer_*):
frames come back from the cache as volumes the render launch samples exactly
as volumes made directly from the same frames; seeks prefetch; the device
budget is kept by least-recently-used eviction; the frame pair around a time
gives the two frames and the blend weight of bh_grmhd_frame_pair."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_streamed_frames_sample_as_direct_volumes(tmp_path):
    from blackhole_amd import grmhd as G
    rng = np.random.default_rng(11)
    frames = [rng.uniform(0.1, 1.0, (8, 6, 12, 4)).astype(np.float32) for _ in range(7)]
    times = [0.0, 1.0, 2.5, 3.0, 4.0, 6.0, 7.5]
    jp, bp = str(tmp_path / "series.json"), str(tmp_path / "series.bin")
    G.write_timeseries(jp, bp, frames, times)
    fb = frames[0].nbytes
    pts = rng.uniform(-45, 45, (300, 3)).astype(np.float32)
    s = G.Streamer(jp, bp, cache_bytes=4 * fb, prefetch=3)
    try:
        print("[stream] created", s.stats(), flush=True)
        assert s.get_frame(0, wait=False) == 0  # nothing is resident yet: a miss that queues frame 0
        for f in (0, 3, 5, 6, 2):
            s.seek(f)
            h = s.get_frame(0, wait=True)
            assert h
            direct = G.Volume(frames[f])
            try:
                got = G.sample(h, 0, 0.0, 1.0, 50.0, pts)
                want = G.sample(direct.handle, 0, 0.0, 1.0, 50.0, pts)
                assert np.array_equal(got, want) and np.abs(want).max() > 0, f
            finally:
                direct.close()
            st = s.stats()
            assert st["resident_bytes"] <= max(4 * fb, 2 * fb) and st["current_frame"] == f
        # prefetch: after a seek the next frames become resident without a request
        s.seek(1)
        assert s.get_frame(0, wait=True)
        nxt = s.get_frame(1, wait=True)
        assert nxt
        st0 = s.stats()
        assert s.get_frame(1) == nxt  # a hit
        assert s.stats()["hits"] == st0["hits"] + 1
        # frame pair around t = 3.5: frames 3 and 4, alpha 0.5
        a, b, al = s.frame_pair(3.5, wait=True)
        assert a and b and abs(al - 0.5) < 1e-6 and s.stats()["current_frame"] == 3
        la, lb = G.Volume(frames[3]), G.Volume(frames[4])
        try:
            got = G.sample(a, b, al, 1.0, 50.0, pts)
            want = G.sample(la.handle, lb.handle, al, 1.0, 50.0, pts)
            assert np.array_equal(got, want)
        finally:
            la.close()
            lb.close()
        st = s.stats()
        print("[stream] final", st, flush=True)
        assert st["resident_frames"] <= 4 and st["resident_bytes"] == st["resident_frames"] * fb
        assert st["hits"] > 0 and st["misses"] > 0 and st["frame_count"] == 7
    finally:
        s.close()


def test_streamer_rejects_bad_inputs(tmp_path):
    from blackhole_amd import grmhd as G
    frames = [np.ones((4, 4, 4, 4), np.float32) for _ in range(3)]
    jp, bp = str(tmp_path / "s.json"), str(tmp_path / "s.bin")
    G.write_timeseries(jp, bp, frames, [0.0, 1.0, 2.0])
    with pytest.raises(RuntimeError, match="budget"):
        G.Streamer(jp, bp, cache_bytes=100)
    with open(bp, "r+b") as f:
        f.truncate(2 * frames[0].nbytes)
    with pytest.raises(RuntimeError, match="beyond the end"):
        G.Streamer(jp, bp, cache_bytes=1 << 20)
