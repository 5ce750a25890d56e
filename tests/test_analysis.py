"""Host-side image measures: closed-form critical curve and shadow size."""
from __future__ import annotations

import math

import numpy as np

from blackhole_amd import abi, analysis, tracer


def test_schwarzschild_shadow_diameter():
    # a = 0: the critical curve is the circle b = 3 sqrt(3)
    assert abs(analysis.shadow_diameter_closed_form(0.0, 1.0) - 6.0 * math.sqrt(3.0)) < 1e-6


def test_critical_curve_points_are_spherical_orbits(oracle):
    """Every curve point maps back (xi = -alpha sin, eta = beta^2 + (alpha^2 - a^2) cos^2)
    to constants with a double root of R: R(r) = R'(r) = 0 for some r."""
    L = oracle.lib()
    a, th = 0.9, math.radians(60.0)
    ca, cb = analysis.critical_curve(a, th, 200)
    rp, rm = analysis.photon_orbits(a)
    for al, be in zip(ca[::17], cb[::17]):
        xi = -al * math.sin(th)
        eta = be * be + (al * al - a * a) * math.cos(th) ** 2
        rs = np.linspace(rp, rm, 4001)
        out = np.zeros(6)
        vals = []
        for r in rs:
            L.orc_potentials(a, xi, eta, r, 1.0, out.ctypes.data_as(oracle._PD))
            vals.append(out[0] / r**4)
        vals = np.array(vals)
        # R touches zero from above on the photon-orbit range (double root)
        assert vals.min() > -1e-6 and abs(vals.min()) < 1e-5


def test_curve_distance_and_shadow_on_synthetic_image():
    cam = tracer.camera(0.0, 60.0, 100.0, 200, 200, 10.0)
    d = analysis.curve_distance(cam)
    A, B = analysis.pixel_grid(cam)
    rho = np.hypot(A, B)
    np.testing.assert_allclose(d, np.abs(rho - 3.0 * math.sqrt(3.0)), atol=0.02)
    img = np.where(rho < 3.0 * math.sqrt(3.0), abi.HIT_HORIZON, abi.HIT_ESCAPED)
    assert abs(analysis.shadow_diameter_pixels(img, cam) - 6.0 * math.sqrt(3.0)) < 0.05


def test_intensity_errors():
    ref = np.array([1.0, 2.0, 0.0, 4.0])
    e = analysis.intensity_errors(ref * (1 + 1e-4), ref)
    assert abs(e["mean_rel"] - 1e-4) < 1e-12 and e["n_pixels"] == 3
