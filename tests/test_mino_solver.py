"""The Mino-time crossing solver (blackhole_amd/csrc/mino.hpp) on the CPU,
compiled by g++ from the same header the GPU kernel includes:

  * Carlson's R_F against scipy's incomplete elliptic integral F(phi|m),
    negative parameters included (the polar motion's parameter is <= 0);
  * the largest radial root against numpy.roots of the same quartic;
  * the polar Mino times against a direct quadrature of du / sqrt(U);
  * the Gauss-Kronrod 3-7 rule of the Newton steps: exact on monomials up
    to degree 11 and not at 12, its embedded 3-point Gauss rule exact up to
    degree 5 and not at 6 (a wrong node or weight breaks one of these);
  * whole rays against the oracle's step-integrated records on the 16 frames
    of the surrogate comparison (spins 0/0.5/0.9/0.998 x 17/40/60/85 deg) at
    low resolution: identical hit classes, crossing radii within 1e-6;
  * the escape direction's pieces -- Carlson's R_C and R_J against scipy's
    elliprc / elliprj, Jacobi sn, cn, dn against scipy's ellipj (m >= 0) and,
    for the polar motion's negative parameters, by F(am(v) | m) = v with
    scipy's ellipkinc and the identities sn^2 + cn^2 = 1, dn^2 = 1 - m sn^2,
    Legendre's Pi against direct quadrature -- and the direction itself
    against the oracle's records at the oracle's own exit radius and, at
    a = 0, against the planar orbit's deflection angle (no integrator).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
from scipy import integrate, special

import oracle_lib as OL
from blackhole_amd import tracer, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_D = C.c_double


@pytest.fixture(scope="module")
def mino(tmp_path_factory):
    so = tmp_path_factory.mktemp("mino") / "libmino_cpu.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    os.path.join(ROOT, "tests", "cpp", "mino_capi.cpp"), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.mino_carlson_rf.restype = _D
    L.mino_carlson_rf.argtypes = [_D] * 3
    L.mino_gk7_monomial.restype = _D
    L.mino_gk7_monomial.argtypes = [C.c_int, _D, _D, C.POINTER(_D)]
    L.mino_radial_roots.argtypes = [_D] * 3 + [C.POINTER(_D)]
    L.mino_polar_times.argtypes = [_D] * 5 + [C.POINTER(_D)]
    L.mino_solve_pixel.argtypes = [_D] * 9 + [C.POINTER(_D), C.c_int, _D, C.POINTER(_D), C.POINTER(C.c_int)]
    for name, n in (("mino_carlson_rc", 2), ("mino_carlson_rj", 4), ("mino_pi3_reduced", 4), ("mino_pi3_complete", 2)):
        getattr(L, name).restype = _D
        getattr(L, name).argtypes = [_D] * n
    L.mino_jacobi.argtypes = [_D, _D, C.POINTER(_D)]
    L.mino_solve_pixel_exit.restype = C.c_int
    L.mino_solve_pixel_exit.argtypes = [_D] * 10 + [C.POINTER(_D), C.POINTER(C.c_int), C.POINTER(_D)]
    return L


def test_carlson_rf_is_the_incomplete_elliptic_integral(mino):
    rng = np.random.default_rng(11)
    for _ in range(500):
        phi = rng.uniform(0.0, math.pi / 2)
        m = -rng.uniform(0.0, 60.0) if rng.random() < 0.7 else rng.uniform(0.0, 0.99)
        s = math.sin(phi)
        got = s * mino.mino_carlson_rf(math.cos(phi) ** 2, 1.0 - m * s * s, 1.0)
        want = special.ellipkinc(phi, m)
        assert abs(got - want) <= 5e-15 * abs(want) + 1e-300
    # complete integral: R_F(0, 1 - m, 1) = K(m)
    for m in (-40.0, -1.0, 0.0, 0.5):
        assert abs(mino.mino_carlson_rf(0.0, 1.0 - m, 1.0) - special.ellipk(m)) <= 5e-15 * special.ellipk(m)


def test_gauss_kronrod_3_7_degrees(mino):
    err = _D()
    for a, b in ((-1.0, 1.0), (0.3, 2.1), (-2.5, -0.4)):
        for d in range(14):
            exact = (b ** (d + 1) - a ** (d + 1)) / (d + 1)
            v = mino.mino_gk7_monomial(d, a, b, C.byref(err))
            scale = max(abs(a), abs(b)) ** d * (b - a)
            k_exact = abs(v - exact) <= 1e-14 * scale
            # the Gauss estimate is v -/+ err: exact iff err vanishes with K7 exact
            g_exact = k_exact and err.value <= 1e-14 * scale
            assert k_exact == (d <= 11 or (d % 2 == 1 and a == -b)), (a, b, d, v, exact)
            assert g_exact == (d <= 5 or (d % 2 == 1 and a == -b)), (a, b, d, err.value)


def test_largest_radial_root_matches_numpy(mino):
    rng = np.random.default_rng(12)
    out = (_D * 2)()
    checked = 0
    for _ in range(3000):
        a, xi, eta = rng.uniform(0, 0.998), rng.uniform(-25, 25), rng.uniform(0.01, 600)
        A, B, Cc = a * a - eta - xi * xi, 2 * (eta + (xi - a) ** 2), -a * a * eta
        roots = np.roots([1.0, 0.0, A, B, Cc])
        real = np.sort(roots[np.abs(roots.imag) <= 1e-9 * np.abs(roots)].real)
        if mino.mino_radial_roots(A, B, Cc, out):
            assert abs(out[1] - real[-1]) <= 1e-12 * abs(real[-1])
            assert abs(((out[1] ** 2 + A) * out[1] + B) * out[1] + Cc) <= 1e-9 * out[1] ** 4
            checked += 1
        else:
            # no r3, r4 pair: the real roots (r1, r2) both lie inside the outer horizon
            assert real[-1] < 1.0 + math.sqrt(1.0 - a * a)
    assert checked > 1000


def test_polar_times_match_quadrature(mino):
    rng = np.random.default_rng(13)
    g = (_D * 2)()
    n = 0
    while n < 150:
        a, th = rng.uniform(0.05, 0.998), math.radians(rng.uniform(5, 85))
        al, be = rng.uniform(-20, 20), rng.uniform(-20, 20)
        uo, so = math.cos(th), math.sin(th)
        xi, eta = -al * so, be * be + (al * al - a * a) * uo * uo
        if eta <= 0.5:
            continue
        assert mino.mino_polar_times(a, xi, eta, uo, so * be, g)
        # U = a^2 (w+ - u^2)(u^2 - w-): the 1/sqrt(w+ - u^2) end handled by the algebraic weight
        b = eta + xi * xi - a * a
        wp = (-b + math.sqrt(b * b + 4 * a * a * eta)) / (2 * a * a)
        wm = -eta / (a * a * wp)
        up = math.sqrt(wp)
        h = lambda u: 1.0 / (a * math.sqrt((up + u) * (u * u - wm)))  # noqa: E731
        K = integrate.quad(h, 0.0, up, weight="alg", wvar=(0.0, -0.5), epsabs=0, epsrel=1e-13)[0]
        Fo = integrate.quad(lambda u: h(u) / math.sqrt(up - u), 0.0, uo, epsabs=0, epsrel=1e-13, limit=200)[0] \
            if uo < 0.999 * up else K - integrate.quad(h, uo, up, weight="alg", wvar=(0.0, -0.5), epsabs=0,
                                                      epsrel=1e-13)[0]
        g0 = Fo if be < 0 else 2 * K - Fo
        assert abs(g[0] - g0) <= 1e-9 * g0 and abs(g[1] - 2 * K) <= 1e-9 * K
        n += 1


def test_rays_match_the_step_integrator(mino):
    w = workloads.c5(24)
    o = w["opts"]
    ref = OL.trace(w["cams"], o, luts=w["luts"], lut_n=w["luts"].shape[1], nthreads=8)
    r = (_D * 3)()
    ints = (C.c_int * 6)()
    n = fails = 0
    worst = 0.0
    for cam in w["cams"]:
        r_stop = o.horizon_factor * tracer.outer_horizon(cam.spin)
        r_in = tracer.isco(cam.spin)
        for j in range(cam.height):
            for i in range(cam.width):
                al = cam.alpha_min + (i + 0.5) * (cam.alpha_max - cam.alpha_min) / cam.width
                be = cam.beta_max - (j + 0.5) * (cam.beta_max - cam.beta_min) / cam.height
                mino.mino_solve_pixel(cam.spin, cam.theta_obs, cam.r_obs, al, be, r_stop, o.r_escape, r_in,
                                      o.disk_r_out, None, 0, 1e-10, r, ints)
                rec = ref[cam.record_offset + j * cam.width + i]
                n += 1
                if not ints[4]:
                    fails += 1
                    continue
                assert ints[2] == rec["hit_class"], (cam.spin, cam.theta_obs, al, be)
                for k in range(min(3, ints[0], rec["n_cross"])):
                    worst = max(worst, abs(r[k] - rec["cross_r"][k]) / rec["cross_r"][k])
    assert fails <= n // 1000
    assert worst <= 1e-6


def test_carlson_rc_rj_match_scipy(mino):
    rng = np.random.default_rng(21)
    worst_c = worst_j = 0.0
    for _ in range(3000):
        x, y = rng.uniform(0.0, 10.0), rng.uniform(0.01, 10.0)
        if rng.random() < 0.3:
            y = x * (1.0 + rng.uniform(-2e-3, 2e-3)) + 1e-9  # the series branch near x = y
        worst_c = max(worst_c, abs(mino.mino_carlson_rc(x, y) / special.elliprc(x, y) - 1.0))
        x, y, z = rng.uniform(0.0, 10.0, 3)
        if rng.random() < 0.2:
            x = 0.0
        p = rng.uniform(0.01, 10.0)
        worst_j = max(worst_j, abs(mino.mino_carlson_rj(x, y, z, p) / special.elliprj(x, y, z, p) - 1.0))
    assert worst_c <= 3e-14 and worst_j <= 1e-14, (worst_c, worst_j)


def test_jacobi_functions(mino):
    rng = np.random.default_rng(22)
    out = (_D * 3)()
    for _ in range(2000):
        m, v = rng.uniform(0.0, 0.999), rng.uniform(-4.0, 4.0)
        mino.mino_jacobi(v, m, out)
        sn, cn, dn, _ = special.ellipj(v, m)
        assert max(abs(out[0] - sn), abs(out[1] - cn), abs(out[2] - dn)) <= 1e-14, (m, v)
    for _ in range(2000):
        m = -rng.uniform(0.0, 60.0)
        K = special.ellipk(m)
        v = rng.uniform(-K, K)
        mino.mino_jacobi(v, m, out)
        am = math.atan2(out[0], out[1])
        assert abs(special.ellipkinc(am, m) - v) <= 1e-13 * max(1.0, K), (m, v)
        assert abs(out[0] ** 2 + out[1] ** 2 - 1.0) <= 5e-14 and abs(out[2] ** 2 - (1.0 - m * out[0] ** 2)) <= 1e-13


def test_legendre_pi_matches_quadrature(mino):
    rng = np.random.default_rng(23)
    for _ in range(300):
        n = rng.uniform(0.0, 0.999)
        m = -rng.uniform(0.0, 40.0) if rng.random() < 0.8 else rng.uniform(0.0, 0.9)
        ph = rng.uniform(-math.pi / 2, math.pi / 2)
        f = lambda t: 1.0 / ((1.0 - n * math.sin(t) ** 2) * math.sqrt(1.0 - m * math.sin(t) ** 2))  # noqa: E731
        want = integrate.quad(f, 0.0, ph, epsabs=0, epsrel=1e-13, limit=200)[0]
        got = mino.mino_pi3_reduced(n, m, math.sin(ph), math.cos(ph))
        assert abs(got - want) <= 1e-13 * abs(want) + 1e-300, (n, m, ph)
        wc = integrate.quad(f, 0.0, math.pi / 2, epsabs=0, epsrel=1e-13, limit=200)[0]
        assert abs(mino.mino_pi3_complete(n, m) - wc) <= 1e-13 * wc


def test_escape_direction_matches_the_step_integrator(mino):
    """u_exit and phi_exit of escaped rays at the oracle's own exit radius (the
    integrator stops at the end of the step that passes r_escape): within the
    integrator's accumulated error."""
    w = workloads.c5(24)
    o = w["opts"]
    ref = OL.trace(w["cams"], o, luts=w["luts"], lut_n=w["luts"].shape[1], nthreads=8)
    r, ints, d = (_D * 3)(), (C.c_int * 6)(), (_D * 2)()
    du, dphi, n, fails = [], [], 0, 0
    for cam in w["cams"]:
        r_stop = o.horizon_factor * tracer.outer_horizon(cam.spin)
        r_in = tracer.isco(cam.spin)
        for j in range(cam.height):
            for i in range(cam.width):
                rec = ref[cam.record_offset + j * cam.width + i]
                if rec["hit_class"] != 3:  # escaped
                    continue
                al = cam.alpha_min + (i + 0.5) * (cam.alpha_max - cam.alpha_min) / cam.width
                be = cam.beta_max - (j + 0.5) * (cam.beta_max - cam.beta_min) / cam.height
                ok = mino.mino_solve_pixel_exit(cam.spin, cam.theta_obs, cam.r_obs, al, be, r_stop, rec["r_exit"],
                                                r_in, o.disk_r_out, 1e-10, r, ints, d)
                n += 1
                if not ok or ints[2] != 3:
                    fails += 1
                    continue
                du.append(abs(d[0] - rec["u_exit"]))
                dphi.append(abs(d[1] - rec["phi_exit"]))
    du, dphi = np.array(du), np.array(dphi)
    print("escape direction vs oracle:", n, "rays; |du| max", du.max(), "|dphi| max", dphi.max(), "mean", dphi.mean())
    assert n > 3000 and fails <= n // 1000
    assert du.max() <= 1e-6 and dphi.max() <= 2e-5 and dphi.mean() <= 1e-6


def test_escape_direction_of_meridional_rays(mino):
    """xi = 0 (the centre column of an odd-width image): the polar term of
    dphi/dl vanishes, as in the integrator; u_exit and phi_exit still match the
    oracle's at its exit radius."""
    w = workloads.c5(25, spins=(0.0, 0.9), inclinations=(30.0, 60.0))
    o = w["opts"]
    ref = OL.trace(w["cams"], o, luts=w["luts"], lut_n=w["luts"].shape[1], nthreads=8)
    r, ints, d = (_D * 3)(), (C.c_int * 6)(), (_D * 2)()
    n = 0
    for cam in w["cams"]:
        i = cam.width // 2
        al = cam.alpha_min + (i + 0.5) * (cam.alpha_max - cam.alpha_min) / cam.width
        assert al == 0.0
        for j in range(cam.height):
            rec = ref[cam.record_offset + j * cam.width + i]
            if rec["hit_class"] != 3:
                continue
            be = cam.beta_max - (j + 0.5) * (cam.beta_max - cam.beta_min) / cam.height
            assert mino.mino_solve_pixel_exit(cam.spin, cam.theta_obs, cam.r_obs, al, be,
                                              o.horizon_factor * tracer.outer_horizon(cam.spin), rec["r_exit"],
                                              tracer.isco(cam.spin), o.disk_r_out, 1e-10, r, ints, d)
            assert abs(d[0] - rec["u_exit"]) <= 1e-6 and abs(d[1] - rec["phi_exit"]) <= 1e-6, (cam.spin, be)
            n += 1
    assert n > 20


def _planar_escape_direction(theta_o, r_o, alpha, beta, r_e):
    """Escape direction of a Schwarzschild (a = 0) ray, independent of the
    integrator and of the Mino-time solver: the orbit lies in the plane through
    the centre spanned by the initial position and the initial angular motion,
    and sweeps the textbook deflection angle psi = int du / sqrt(1/b^2 - u^2 +
    2 u^3) (u = 1/r, b^2 = eta + xi^2) from r_o in to the periapsis and out to
    r_e.  Returns (cos theta, phi) at r_e, phi unwrapped from phi_o = 0."""
    so, co = math.sin(theta_o), math.cos(theta_o)
    xi = -alpha * so
    b2 = beta * beta + xi * xi / (so * so)
    f = lambda u: 1.0 / b2 - u * u + 2.0 * u ** 3
    from scipy.optimize import brentq
    up = brentq(f, 1.0 / r_o, 1.0 / 3.0, xtol=1e-17, rtol=1e-15)
    g = lambda s: 2.0 * s / math.sqrt(f(up - s * s))  # u = up - s^2 removes the root
    psi = sum(integrate.quad(g, 0.0, math.sqrt(up - 1.0 / r), epsabs=1e-15, epsrel=1e-12, limit=400)[0]
              for r in (r_o, r_e))
    rhat = np.array([so, 0.0, co])
    # initial angular motion: d theta/d lambda = -beta (du/d lambda = sin theta_o beta),
    # sin theta d phi/d lambda = xi / sin theta_o
    t = -beta * np.array([co, 0.0, -so]) + (xi / so) * np.array([0.0, 1.0, 0.0])
    t /= np.linalg.norm(t)
    ell_z = np.cross(rhat, t)[2]
    # |d phi / d psi| <= 1 / |ell_z| (closest approach to the pole): sample finely enough to unwrap
    n = int(max(2e4, 50.0 * psi / max(abs(ell_z), 1e-12)))
    ps = np.linspace(0.0, psi, n + 1)
    x = np.cos(ps)[:, None] * rhat + np.sin(ps)[:, None] * t
    phi = np.unwrap(np.arctan2(x[:, 1], x[:, 0]))
    return float(x[-1, 2]), float(phi[-1])


def test_escape_direction_against_the_planar_schwarzschild_orbit(mino):
    """a = 0: the solver's u_exit, phi_exit against the planar orbit's closed
    form, including rays that pass within 5e-4 rad of the pole, where
    d phi / d lambda = xi / sin^2 theta is a narrow spike.  (The step
    integrator's phi_exit is 0.17 rad off for such rays: its embedded error
    estimate does not see a spike that falls between its stages.)"""
    r, ints, d = (_D * 3)(), (C.c_int * 6)(), (_D * 2)()
    r_stop = 1.01 * 2.0
    n = 0
    for th_deg in (17.0, 60.0):
        th = math.radians(th_deg)
        for al, be in ((-0.009765625, 5.224609375), (0.009765625, 5.224609375), (0.01, -7.5),
                       (6.0, 2.0), (-3.0, -5.0), (7.0, 0.0), (0.5, 12.0)):
            for r_e in (200.0, 500.0):
                # no disk (r_out < r_in): every crossing passes through
                assert mino.mino_solve_pixel_exit(0.0, th, 100.0, al, be, r_stop, r_e, 6.0, 0.0, 1e-10,
                                                  r, ints, d)
                assert ints[2] == 3
                u, phi = _planar_escape_direction(th, 100.0, al, be, r_e)
                assert abs(d[0] - u) <= 1e-9 and abs(d[1] - phi) <= 1e-8, (th_deg, al, be, r_e, d[0], u, d[1], phi)
                n += 1
    assert n == 28
