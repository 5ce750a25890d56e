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


def _inside_curve_image(cam, scale=1.0, covered=None):
    """Hit-class image that is HORIZON strictly inside the closed-form
    critical curve (scaled about its centroid) and ESCAPED outside; polar
    angles in `covered` (a (lo, hi) range in rad) are DISK (a disk in front)."""
    ca, cb = analysis.critical_curve(cam.spin, cam.theta_obs, 20000)
    c = analysis.curve_centre(cam.spin, cam.theta_obs)
    psi_c = np.mod(np.arctan2(cb - c[1], ca - c[0]), 2 * math.pi)
    rho_c = np.hypot(ca - c[0], cb - c[1]) * scale
    o = np.argsort(psi_c)
    A, B = analysis.pixel_grid(cam)
    psi = np.mod(np.arctan2(B - c[1], A - c[0]), 2 * math.pi)
    rho = np.hypot(A - c[0], B - c[1])
    r_at = np.interp(psi, psi_c[o], rho_c[o], period=2 * math.pi)
    img = np.where(rho < r_at, abi.HIT_HORIZON, abi.HIT_ESCAPED).astype(np.uint8)
    if covered is not None:
        img[(psi > covered[0]) & (psi < covered[1]) & (rho > 0.5 * r_at)] = abi.HIT_DISK
    return img


def test_ring_diameter_estimator_on_synthetic_images():
    """The shadow-edge ring estimator: on an image that is captured exactly
    inside the Bardeen curve it reproduces the curve's diameter to within a
    pixel; it sees a 1 % larger ring as 1 %; bins where a disk covers the
    curve are left out of the ring; with the whole curve covered it falls
    back to the shadow outline."""
    for a, th in ((0.0, 17.0), (0.9, 60.0), (0.998, 85.0)):
        cam = tracer.camera(a, th, 100.0, 1024, 1024, 10.0)
        px = 20.0 / 1024
        img = _inside_curve_image(cam)
        e = analysis.ring_diameter_errors(img, img, cam)
        assert e["measured_on"] == "ring" and e["ring_bins"] == 64 and e["rel_err"] == 0.0
        assert abs(e["diameter_ref"] - e["curve_diameter"]) <= 2 * px, e
        big = _inside_curve_image(cam, 1.01)
        e = analysis.ring_diameter_errors(big, img, cam)
        assert abs(e["rel_err"] - 0.01) <= 2 * px / e["diameter_ref"], e
        part = _inside_curve_image(cam, covered=(0.5, 2.0))
        e = analysis.ring_diameter_errors(part, part, cam)
        assert 40 <= e["ring_bins"] < 64 and abs(e["diameter_ref"] - e["curve_diameter"]) <= 2 * px, e
        hidden = _inside_curve_image(cam, covered=(-1.0, 7.0))
        e = analysis.ring_diameter_errors(hidden, hidden, cam)
        assert e["ring_bins"] == 0 and e["measured_on"] == "shadow"
        assert abs(e["diameter"] - 0.5 * analysis.diameter_from_radii(analysis.curve_radii(cam))) <= 4 * px
