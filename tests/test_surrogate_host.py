"""Host-side pieces of the surrogate pipeline (no GPU): the critical-curve
polar tables that define the fallback band, against the brute-force distance
of every pixel to the Bardeen curve."""
from __future__ import annotations

import math

import numpy as np

from blackhole_amd import analysis, sur_render, tracer


def _band_flags(table: np.ndarray, cam, band: float) -> np.ndarray:
    nb = sur_render.S.RING_BINS
    A, B = analysis.pixel_grid(cam)
    da, db = A.reshape(-1) - table[0], B.reshape(-1) - table[1]
    psi = np.arctan2(db, da)
    psi = np.where(psi < 0, psi + 2 * math.pi, psi)
    b = np.minimum((psi * (nb / (2 * math.pi))).astype(int), nb - 1)
    rho = np.hypot(da, db)
    lo, hi, obl = table[2 + b], table[2 + nb + b], table[2 + 2 * nb + b]
    d = np.maximum(np.maximum(lo - rho, rho - hi), 0.0)
    return d < band * obl


def test_ring_tables_cover_the_distance_band():
    """Every pixel within the band's distance of the critical curve is flagged
    (no misses), and the flagged set is at most 15 % larger than the band."""
    cams = [tracer.camera(a, th, 100.0, 160, 160, 20.0)
            for a, th in ((0.0, 40.0), (0.5, 17.0), (0.9, 60.0), (0.998, 85.0), (0.998, 17.0))]
    T = sur_render.ring_tables(cams)
    for band in (0.1, 0.25, 0.5):
        for k, cam in enumerate(cams):
            d_true = analysis.curve_distance(cam).reshape(-1)
            flag = _band_flags(T[k], cam, band)
            inside = d_true < band
            assert not (inside & ~flag).any(), (cam.spin, cam.theta_obs, band)
            assert flag.sum() <= 1.15 * inside.sum() + 20


def test_ring_table_of_schwarzschild_is_the_circle():
    T = sur_render.ring_tables([tracer.camera(0.0, 45.0, 100.0, 8, 8, 20.0)])[0]
    nb = sur_render.S.RING_BINS
    b = 3.0 * math.sqrt(3.0)
    assert abs(T[0]) < 1e-6 and T[1] == 0.0
    assert np.allclose(T[2:2 + nb], b, atol=1e-5) and np.allclose(T[2 + nb:2 + 2 * nb], b, atol=1e-5)
    assert np.allclose(T[2 + 2 * nb:], 1.0, atol=1e-3)
