"""Pins the FP64 CPU oracle before the GPU results are checked against it.

Sources, in order of preference:
  * golden fixtures of the reference (tests/golden/*.csv, copied from the
    reference's assets/luts and assets/validation),
  * known answers of the reference's own tests (tests/rte_integrator_test.cpp,
    src/physics/physics_test.cpp, tests/verified_physics_test.cpp),
  * the reference's own functions compiled from its sources into
    oracle/_ref/libref.so (oracle/build_ref.sh; skipped where absent),
  * closed forms and invariants chosen so that a dropped term, a wrong sign
    or a transposed operand fails (turning points, capture threshold,
    deflection integral, spherical photon orbits, convergence order).
"""
from __future__ import annotations

import ctypes as C
import csv
import json
import math
import os
import subprocess

import numpy as np
import pytest

from blackhole_amd import abi, flops

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_PD = C.POINTER(C.c_double)
_D = C.c_double


def arr(*v):
    return np.array(v, dtype=np.float64)


def P(a):
    return a.ctypes.data_as(_PD)


def read_csv(name):
    with open(os.path.join(GOLD, name)) as f:
        rows = list(csv.reader(f))
    return rows[0], np.array([[float(x) for x in r] for r in rows[1:]])


@pytest.fixture(scope="module")
def L(oracle):
    return oracle.lib()


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libref.so not built (needs /root/reference)")
    R = C.CDLL(REF_LIB)
    R.ref_kerr_potentials.argtypes = [_D] * 6 + [_PD]
    R.ref_carter_constant.argtypes = [_D] * 5
    R.ref_carter_constant.restype = _D
    R.ref_critical_params.argtypes = [_D, _D, _PD]
    R.ref_photon_orbit.argtypes = [_D, C.c_int]
    R.ref_photon_orbit.restype = _D
    R.ref_isco_geom.argtypes = [_D]
    R.ref_isco_geom.restype = _D
    R.ref_kerr_isco.argtypes = [_D, C.c_int]
    R.ref_kerr_isco.restype = _D
    R.ref_outer_horizon.argtypes = [_D]
    R.ref_outer_horizon.restype = _D
    R.ref_disk_g.argtypes = [_D] * 4
    R.ref_disk_g.restype = _D
    R.ref_rte_step.argtypes = [_PD, _D, _D, _D]
    R.ref_planck.argtypes = [_D, _D]
    R.ref_planck.restype = _D
    R.ref_synch_j.argtypes = [_D] * 4
    R.ref_synch_j.restype = _D
    R.ref_synch_alpha.argtypes = [_D] * 4
    R.ref_synch_alpha.restype = _D
    R.ref_kerr_step_mino.argtypes = [_D, _D, _D, _D, _PD, _D]
    R.ref_disk_flux.argtypes = [_D] * 4
    R.ref_disk_flux.restype = _D
    for name in ("ref_gr_transform_intensity", "ref_gr_transform_emission", "ref_gr_transform_absorption"):
        getattr(R, name).argtypes = [_D, _D]
        getattr(R, name).restype = _D
    R.ref_kerr_gcon.argtypes = [_D, _D, _D, _D, _PD]
    return R


# ---------------------------------------------------------------------------
# Golden fixtures of the reference
# ---------------------------------------------------------------------------

def test_emissivity_lut_matches_reference_fixture(oracle):
    """assets/luts/emissivity_lut.csv: a = 0, NT flux over [r_isco, 4 r_isco]
    (lut_meta.json: r_in 3 r_s, r_out 12 r_s), normalised to its peak."""
    _, tab = read_csv("emissivity_lut.csv")
    meta = json.load(open(os.path.join(GOLD, "lut_meta.json")))
    assert meta["spin"] == 0.0 and meta["size"] == 256
    r_in = 2.0 * meta["r_in_over_rs"]
    r_out = 2.0 * meta["r_out_over_rs"]
    got = oracle.disk_lut(0.0, r_in, r_out, 256)
    np.testing.assert_allclose(tab[:, 0], np.arange(256) / 255.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(got, tab[:, 1], rtol=1e-12, atol=1e-15)


def test_spin_radii_fixture(L):
    """assets/luts/spin_radii_lut.csv + assets/validation/spin_radii_curve.csv:
    ISCO and prograde photon orbit in units of r_s = 2M.  Rows with a < 0 are
    not used: the fixture's r_ph column is wrong there (0.584 r_s at
    a = -0.99, a retrograde orbit must lie beyond 3 M)."""
    for name in ("spin_radii_lut.csv", "spin_radii_curve.csv"):
        _, tab = read_csv(name)
        for a, risco_rs, rph_rs in tab:
            if a < 0:
                continue
            assert abs(L.orc_isco(a) / 2.0 - risco_rs) <= 1e-12 * risco_rs
            assert abs(L.orc_photon_orbit(a, 1) / 2.0 - rph_rs) <= 1e-12 * rph_rs


def test_redshift_fixture_static_emitter(L):
    """assets/luts/redshift_lut.csv / validation/redshift_curve.csv:
    z = 1/sqrt(1 - r_s/r) - 1 (a = 0, equator).  The oracle's fluid redshift
    with zero angular momentum (ell = 0) at a = 0 is a static emitter,
    g = sqrt(1 - 2/r) = 1/(1 + z), for any photon xi."""
    _, curve = read_csv("redshift_curve.csv")
    for r_rs, z in curve:
        g = L.orc_fluid_g(2.0 * r_rs, 0.0, 0.0, 3.7, 0.0)
        assert abs((1.0 / g - 1.0) - z) <= 1e-12 * max(z, 1e-3)
    _, lut = read_csv("redshift_lut.csv")
    for u, z in lut:
        r = 6.0 + u * 18.0
        g = L.orc_fluid_g(r, 0.0, 0.0, -2.0, 0.0)
        assert abs((1.0 / g - 1.0) - z) <= 1e-11 * z


# ---------------------------------------------------------------------------
# Known answers of the reference's tests
# ---------------------------------------------------------------------------

def test_kerr_closed_form_known_answers(L):
    """src/physics/physics_test.cpp:477-641 / tests/novikov_thorne_test.cpp:37,120:
    r_+ = 2 at a = 0; ISCO 6 (a = 0), 1 (a -> 1), ~1.24 at a = 0.998;
    photon orbit 3 (a = 0), 1 (a = 1, prograde), 4 (a = 1, retrograde)."""
    assert L.orc_outer_horizon(0.0) == 2.0
    assert abs(L.orc_isco(0.0) - 6.0) < 1e-12
    assert abs(L.orc_isco(0.998) - 1.2370) < 1e-3
    assert abs(L.orc_photon_orbit(0.0, 1) - 3.0) < 1e-12
    assert abs(L.orc_photon_orbit(1.0, 1) - 1.0) < 1e-12
    assert abs(L.orc_photon_orbit(1.0, 0) - 4.0) < 1e-12


def _rte(L, I0, j, alpha, ds):
    s = arr(I0, 0.0, 0.0)
    L.orc_rte_step(P(s), j, alpha, ds)
    return s


def test_rte_known_answers(L):
    """tests/rte_integrator_test.cpp:112-419: pure emission I = j L (1e-12),
    pure absorption I0 e^{-alpha L} (1e-10), thick slab -> S (1e-6),
    optically thin alpha L (1e-10), Kirchhoff j/alpha = B_nu (1e-10),
    Rayleigh-Jeans (1e-4) and Wien (1e-3) limits of B_nu."""
    s = _rte(L, 0.0, 2.5e-20, 0.0, 3.0e12)
    assert abs(s[0] - 2.5e-20 * 3.0e12) <= 1e-12 * s[0]
    s = _rte(L, 7.0, 0.0, 2e-12, 4e11)
    assert abs(s[0] - 7.0 * math.exp(-0.8)) <= 1e-10 * s[0]
    assert abs(s[1] - 0.8) < 1e-15
    s = _rte(L, 0.0, 5e-10, 1e-9, 1e12)  # tau = 1000 -> S = j/alpha
    assert abs(s[0] - 0.5) <= 1e-6 * 0.5
    s = _rte(L, 0.0, 3e-10, 1e-20, 1e10)  # tau = 1e-10 -> I ~ j L
    assert abs(s[0] - 3.0) <= 1e-9 * 3.0
    h, k, c = 2 * math.pi * 1.054571817e-27, 1.380649e-16, 2.99792458e10
    for nu, T in ((1e9, 1e4), (230e9, 1.7e11)):  # h nu << k T
        rj = 2 * nu * nu * k * T / c**2
        assert abs(L.orc_planck(nu, T) - rj) <= 1e-4 * rj
    nu, T = 1e16, 1e4  # h nu >> k T
    wien = 2 * h * nu**3 / c**2 * math.exp(-h * nu / (k * T))
    assert abs(L.orc_planck(nu, T) - wien) <= 1e-3 * wien


# ---------------------------------------------------------------------------
# The reference's own functions, compiled from its sources
# ---------------------------------------------------------------------------

def test_potentials_match_reference_kerr_cpp(L, ref):
    """R and R' equal kerrPotentials (src/physics/kerr.cpp:69-97).  Its
    Theta(theta) = Q + a^2 cos^2 - L^2/sin^2 differs from the oracle's
    Q + a^2 cos^2 - L^2 cot^2 by exactly L^2: the reference's own Carter
    constant (verified/geodesic.hpp:178-188) fixes the oracle's form, checked
    by recovering Q from p_theta^2 = Theta_oracle."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.uniform(0, 0.999)
        xi = rng.uniform(-8, 8)
        eta = rng.uniform(0, 40)
        r = rng.uniform(1.5, 80)
        th = rng.uniform(0.2, 2.9)
        out = np.zeros(6)
        L.orc_potentials(a, xi, eta, r, th, P(out))
        rr = np.zeros(4)
        ref.ref_kerr_potentials(r, th, a, 1.0, xi, eta, P(rr))
        assert abs(out[0] - rr[0]) <= 1e-12 * max(abs(rr[0]), r**4)
        assert abs(out[1] - rr[1]) <= 1e-12 * max(abs(rr[1]), r**3)
        assert abs((out[4] - rr[2]) - xi * xi) <= 1e-10 * max(1.0, xi * xi)
        # U(u) = sin^2(theta) Theta(theta)
        assert abs(out[2] - math.sin(th) ** 2 * out[4]) <= 1e-11 * max(1.0, abs(out[2]))
        if out[4] > 0:
            q = ref.ref_carter_constant(th, a, 1.0, xi, math.sqrt(out[4]))
            assert abs(q - eta) <= 1e-10 * max(1.0, eta)


def test_phi_rate_matches_reference_mino_derivative(L, ref):
    """dphi/dlambda and |dr/dlambda| of kerrMinoDerivatives (kerr.cpp:40-66),
    read off one tiny kerrStepMino step, equal the oracle's rates x r (the
    oracle's step variable is ds = r dlambda)."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = rng.uniform(0, 0.99)
        xi = rng.uniform(-5, 5)
        eta = rng.uniform(1, 30)
        r = rng.uniform(8, 60)
        th = rng.uniform(0.6, 2.5)
        out = np.zeros(6)
        L.orc_potentials(a, xi, eta, r, th, P(out))
        if out[0] <= 0 or out[4] <= 0:
            continue
        u = math.cos(th)
        y = arr(r, u, 0.0, -math.sqrt(out[0]), -math.sin(th) * math.sqrt(out[4]))
        f = np.zeros(5)
        L.orc_rhs(a, xi, eta, P(y), P(f))
        dl = 1e-7
        st = arr(r, th, 0.0, 0.0, -1.0, 1.0)
        ref.ref_kerr_step_mino(a, 1.0, xi, eta, P(st), dl)
        dphi_ref = st[2] / dl
        # second-order Taylor of r(lambda): r'' = R'/2
        dr_pred = f[0] * r * dl + 0.25 * out[1] * dl * dl
        assert abs(f[2] * r - dphi_ref) <= 1e-6 * max(1.0, abs(dphi_ref))
        assert abs(dr_pred - (st[0] - r)) <= 1e-9 * abs(st[0] - r)


def test_critical_curve_matches_reference(L, ref):
    """criticalImpactParams (analytic_kerr_geodesic.h:410-448) and, as a
    closed-form check on both, R(r_ph) = R'(r_ph) = 0 (spherical orbit)."""
    for a in (0.3, 0.5, 0.9, 0.94, 0.998):
        rp, rm = L.orc_photon_orbit(a, 1), L.orc_photon_orbit(a, 0)
        for r in np.linspace(rp + 1e-3, rm - 1e-3, 9):
            mine, theirs = np.zeros(2), np.zeros(2)
            L.orc_critical_params(r, a, P(mine))
            ref.ref_critical_params(r, a, P(theirs))
            np.testing.assert_allclose(mine, theirs, rtol=1e-13, atol=1e-12)
            out = np.zeros(6)
            L.orc_potentials(a, mine[0], mine[1], r, 1.0, P(out))
            assert abs(out[0]) <= 1e-10 * r**4 and abs(out[1]) <= 1e-10 * r**3


def test_closed_forms_match_reference(L, ref):
    for a in (0.0, 0.3, 0.9, 0.94, 0.998):
        assert abs(L.orc_isco(a) - ref.ref_isco_geom(a)) <= 1e-13
        assert abs(L.orc_isco(a) - ref.ref_kerr_isco(a, 1)) <= 1e-12
        assert abs(L.orc_outer_horizon(a) - ref.ref_outer_horizon(a)) <= 1e-12
        for pro in (0, 1):
            assert L.orc_photon_orbit(a, pro) == ref.ref_photon_orbit(a, pro)


def test_disk_g_matches_reference_cunningham_at_a0(L, ref):
    """kerrDiskGFactor (iron_kline.h:90-109) approximates the photon's
    xi by r sin(phi) sin(i) and drops u^t's r^1.5/(r^1.5 + a) factor; with
    that xi the two agree exactly at a = 0, and at a > 0 differ by that factor."""
    for r in (6.5, 8.0, 15.0):
        for phi in (0.3, 1.2, 2.8, 4.0):
            for inc in (0.2, 1.0, 1.5):
                xi = r * math.sin(phi) * math.sin(inc)
                mine = L.orc_disk_g(r, 0.0, xi)
                theirs = ref.ref_disk_g(r, phi, 0.0, inc)
                assert abs(mine - theirs) <= 1e-13 * theirs
                a = 0.9
                mine = L.orc_disk_g(r, a, xi)
                theirs = ref.ref_disk_g(r, phi, a, inc)
                r32 = r * math.sqrt(r)
                assert abs(mine - theirs * r32 / (r32 + a)) <= 1e-12 * mine


def test_rte_and_synchrotron_match_reference(L, ref):
    """rteStep, planckFunction and the thermal synchrotron j, alpha of
    rte_integrator.h:138-357.  The reference compiled here has no Boost, so
    its K_2(1/Theta) is the documented fallback 2 Theta^2 -- the oracle's
    k2 ratio 1; the oracle uses the exact K_2 ratio from its table."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        nu = 10 ** rng.uniform(9, 13)
        B = 10 ** rng.uniform(-1, 2)
        ne = 10 ** rng.uniform(-2, 7)
        th = 10 ** rng.uniform(-0.5, 2.5)
        j_ref = ref.ref_synch_j(nu, B, ne, th)
        j = L.orc_synch_j(nu, B, ne, th, 1.0)
        assert abs(j - j_ref) <= 1e-12 * j_ref
        T = th * 9.1093837015e-28 * 2.99792458e10**2 / 1.380649e-16
        assert abs(L.orc_planck(nu, T) - ref.ref_planck(nu, T)) <= 1e-13 * ref.ref_planck(nu, T)
        a_ref = ref.ref_synch_alpha(nu, B, ne, th)
        assert abs(j / L.orc_planck(nu, T) - a_ref) <= 1e-12 * a_ref
    for I0, j, alpha, ds in ((0.0, 1.0, 2.0, 3.0), (5.0, 1e-3, 1e-2, 30.0), (2.0, 0.5, 0.0, 4.0)):
        a1, a2 = arr(I0, 0.0, 0.0), arr(I0, 0.0, 0.0)
        L.orc_rte_step(P(a1), j, alpha, ds)
        ref.ref_rte_step(P(a2), j, alpha, ds)
        np.testing.assert_allclose(a1, a2, rtol=1e-14)
    # Below tau = 1e-4 the reference switches to a 2nd-order Taylor form
    # (rte_integrator.h:154-159; stated accuracy < 1e-8 relative).
    a1, a2 = arr(1.0, 0.0, 0.0), arr(1.0, 0.0, 0.0)
    L.orc_rte_step(P(a1), 3e-5, 1e-5, 2.0)
    ref.ref_rte_step(P(a2), 3e-5, 1e-5, 2.0)
    np.testing.assert_allclose(a1, a2, rtol=1e-8)


def test_render_abi_layout_matches_reference_header():
    exe = os.path.join(ROOT, "oracle", "_ref", "abi_layout")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/abi_layout not built")
    lay = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    golden = json.load(open(os.path.join(GOLD, "bh_launch_params_layout.json")))
    assert lay == golden


def test_render_abi_layout_golden_matches_mirror():
    golden = json.load(open(os.path.join(GOLD, "bh_launch_params_layout.json")))
    assert C.sizeof(abi.BH_LaunchParams) == golden["__sizeof__"] == 336
    for name, *_ in abi.BH_LaunchParams._fields_:
        assert getattr(abi.BH_LaunchParams, name).offset == golden[name], name


# ---------------------------------------------------------------------------
# Closed forms and invariants of the integrator
# ---------------------------------------------------------------------------

def _opts(**kw):
    o = abi.BH_TraceOptions()
    o.integrator = 0
    o.max_steps = 20000
    o.max_attempts = 200000
    o.max_crossings = 3
    o.h_fixed = 0.005
    o.h_init = 0.01
    o.h_max = 0.05
    o.h_vol = 0.01
    o.rtol = 1e-11
    o.atol = 1e-13
    o.horizon_factor = 1.01
    o.r_escape = 2000.0
    o.disk_r_out = 20.0
    o.nu_obs = 230e9
    o.disk_T_max = 1e7
    o.length_unit_cm = 5.9e11
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _one_ray(oracle, a, theta_deg, r_obs, alpha, beta, **kw):
    cam = abi.BH_Camera(a, math.radians(theta_deg), r_obs, alpha - 1e-3, alpha + 1e-3,
                        beta - 1e-3, beta + 1e-3, 1, 1, 0)
    return oracle.trace([cam], _opts(**kw))[0]


@pytest.mark.parametrize("integ", [0, 1])
def test_schwarzschild_capture_threshold(oracle, integ):
    """a = 0: captured iff b < 3 sqrt(3) M (src/physics/geodesics.cpp:26-35)."""
    bc = 3.0 * math.sqrt(3.0)
    for b, cls in ((bc - 0.01, abi.HIT_HORIZON), (bc + 0.01, abi.HIT_ESCAPED)):
        rec = _one_ray(oracle, 0.0, 90.0, 1000.0, b, 0.0, integrator=integ)
        assert rec["hit_class"] == cls, (b, rec)


def test_schwarzschild_deflection_integral(oracle):
    """Equatorial a = 0 ray: total phi swept from r_obs in to r_min and out to
    r_exit equals the quadrature int dr / (r^2 sqrt(1/b^2 - (1 - 2/r)/r^2))."""
    from scipy.integrate import quad
    from scipy.optimize import brentq
    for b in (7.0, 12.0, 40.0):
        rec = _one_ray(oracle, 0.0, 90.0, 500.0, b, 0.0, integrator=1)
        assert rec["hit_class"] == abi.HIT_ESCAPED
        rmin = brentq(lambda r: r**3 - b * b * (r - 2.0), 3.0, b + 1)
        f = lambda x, rr: 1.0 / math.sqrt(1.0 / b**2 - (1 - 2 / rr) / rr**2) / rr**2
        # substitute r = rmin + x^2 to remove the endpoint singularity
        g = lambda x: 2 * x * f(x, rmin + x * x)
        phi = quad(g, 0, math.sqrt(500.0 - rmin), epsabs=1e-12, epsrel=1e-12, limit=400)[0]
        phi += quad(g, 0, math.sqrt(rec["r_exit"] - rmin), epsabs=1e-12, epsrel=1e-12, limit=400)[0]
        assert abs(abs(rec["phi_exit"]) - phi) <= 1e-8 * phi, (b, rec["phi_exit"], phi)


def test_kerr_equatorial_turning_point(L):
    """a = 0.9, equatorial ray with b = 6 (xi = 6, eta = 0) coming in from
    r = 12: its radial turning point is the largest root of R(r), r = 4.972845
    (the value at which the reference's first-order stepper stalls), and the
    second-order form turns there instead of stalling."""
    a, xi, eta = 0.9, 6.0, 0.0
    out = np.zeros(6)
    L.orc_potentials(a, xi, eta, 12.0, math.pi / 2, P(out))
    y = arr(12.0, 0.0, 0.0, -math.sqrt(out[0]), 0.0)
    yn = np.zeros(5)
    r_tp = None
    for _ in range(4000):
        L.orc_rk4_step(a, xi, eta, P(y), 0.002, P(yn))
        if y[3] < 0 <= yn[3]:
            # v_r changed sign: one Newton step on R(r) = v_r^2 from the state
            # after the turn, r_tp = r - v_r^2 / R'(r)
            L.orc_potentials(a, xi, eta, yn[0], math.pi / 2, P(out))
            r_tp = yn[0] - yn[3] ** 2 / out[1]
        y[:] = yn
        if y[3] > 0 and y[0] > 20:
            break
    from scipy.optimize import brentq
    Rf = lambda r: (r * r + a * a - a * xi) ** 2 - (r * r - 2 * r + a * a) * (eta + (xi - a) ** 2)
    root = brentq(Rf, 4.0, 6.0, xtol=1e-14)
    assert abs(root - 4.972845) < 1e-6
    assert r_tp is not None and abs(r_tp - root) < 1e-8, (r_tp, root)
    assert y[3] > 0  # turned around (no stall)


def test_rk4_and_dp45_convergence_order(L):
    """Local error of one step on the Kerr system scales as h^5 (RK4) and
    h^6 (DP45 5th-order solution) against a reference from 64 RK4 substeps."""
    a, xi, eta = 0.94, 2.5, 14.0
    out = np.zeros(6)
    L.orc_potentials(a, xi, eta, 7.0, 1.1, P(out))
    y0 = arr(7.0, math.cos(1.1), 0.3, -math.sqrt(out[0]), math.sin(1.1) * math.sqrt(out[4]))

    def rk4_n(h, n):
        y = y0.copy()
        yn = np.zeros(5)
        for _ in range(n):
            L.orc_rk4_step(a, xi, eta, P(y), h / n, P(yn))
            y[:] = yn
        return y

    errs_rk4, errs_dp = [], []
    for h in (0.08, 0.04):
        exact = rk4_n(h, 256)
        e1 = np.abs(rk4_n(h, 1) - exact).max()
        y5 = np.zeros(5)
        L.orc_dp45_attempt(a, xi, eta, P(y0), h, 1e-8, 1e-10, P(y5))
        e2 = np.abs(y5 - exact).max()
        errs_rk4.append(e1)
        errs_dp.append(e2)
    assert 24 < errs_rk4[0] / errs_rk4[1] < 40      # 2^5 = 32
    assert 40 < errs_dp[0] / errs_dp[1] < 90        # 2^6 = 64


def test_first_integrals_conserved(oracle):
    """Along full traces, v_r^2 - R(r) and v_u^2 - U(u) stay at round-off
    relative to their scales (null condition and Carter constant)."""
    cam = abi.BH_Camera(0.9, math.radians(60), 100.0, -15, 15, -15, 15, 24, 24, 0)
    lut = oracle.disk_lut(0.9, 2.3208830417618866, 20.0)
    rec = oracle.trace([cam], _opts(integrator=0, h_fixed=0.01, r_escape=200.0, disk_enabled=1),
                       luts=lut, lut_n=256)
    assert rec["drift_null"].max() < 1e-8
    assert rec["drift_carter"].max() < 1e-7


def test_pow_m0p2_and_controller(L):
    for x in np.concatenate([10.0 ** np.linspace(-12, 12, 97), [0.5, 1.0, 2.0, 3.0]]):
        assert abs(L.orc_pow_m0p2(x) - x ** -0.2) <= 2e-6 * x ** -0.2
    assert L.orc_dp45_factor(1e-30, 1, 0) == 5.0
    assert L.orc_dp45_factor(1e-30, 1, 1) == 1.0
    assert abs(L.orc_dp45_factor(2.0, 0, 0) - 0.9 * 2.0 ** -0.2) < 2e-6
    assert L.orc_dp45_factor(1e9, 0, 0) == 0.2


def test_hermite_crossing_exact_for_cubic(L):
    """A cubic u(s) and r(s) are reproduced exactly by the Hermite form, so
    the refined crossing radius equals r at the cubic's root."""
    h = 0.3
    uu = lambda s: 0.4 - 2.0 * s + 1.5 * s * s - 0.7 * s ** 3
    du = lambda s: -2.0 + 3.0 * s - 2.1 * s * s
    rr = lambda s: 9.0 - 3.0 * s + 0.5 * s ** 3
    dr = lambda s: -3.0 + 1.5 * s * s
    y0 = arr(rr(0), uu(0), 0, 0, 0)
    y1 = arr(rr(h), uu(h), 0, 0, 0)
    f0 = arr(dr(0), du(0), 0, 0, 0)
    f1 = arr(dr(h), du(h), 0, 0, 0)
    from scipy.optimize import brentq
    s0 = brentq(uu, 0, h)
    assert abs(L.orc_crossing_radius(P(y0), P(f0), P(y1), P(f1), h) - rr(s0)) < 1e-13


def test_bessel_k_quadrature(L):
    """K_nu by quadrature vs the closed form K_1/2(z) = sqrt(pi/2z) e^-z and
    scipy.special.kv; large-Theta limit K_2(1/Theta) -> 2 Theta^2."""
    from scipy.special import kv
    for z in (1e-3, 0.05, 0.5, 2.0, 20.0, 200.0):
        assert abs(L.orc_bessel_k(0.5, z) - math.sqrt(math.pi / (2 * z)) * math.exp(-z)) <= 1e-12 * L.orc_bessel_k(0.5, z)
        assert abs(L.orc_bessel_k(2.0, z) - kv(2, z)) <= 1e-12 * kv(2, z)
    assert abs(L.orc_bessel_k(2.0, 1e-3) / (2e6) - 1.0) < 1e-6


def test_volume_fetch_trilinear_bruteforce(L):
    """Cell-centred trilinear sampling (four cell fields) with r clamp, theta
    clamp and phi wrap, checked against a direct numpy evaluation on a random
    6 x 5 x 4 grid."""
    rng = np.random.default_rng(4)
    nr, nt, npz = 6, 5, 4
    cells = rng.uniform(0, 1, (npz, nt, nr, 4)).astype(np.float32)
    vol = abi.BH_Volume(None, nr, nt, npz, 0, 2.0, 14.0, 0, 0, 0, 0)

    def brute(r, u, phi):
        th = math.acos(u)
        xs = [(r - 2.0) / 12.0 * nr - 0.5, th / math.pi * nt - 0.5, phi / (2 * math.pi) * npz - 0.5]
        out = np.zeros(4)
        for f in range(4):
            acc = 0.0
            i0 = [math.floor(x) for x in xs]
            fr = [x - math.floor(x) for x in xs]
            for dr_ in (0, 1):
                for dt_ in (0, 1):
                    for dp_ in (0, 1):
                        ir = min(max(i0[0] + dr_, 0), nr - 1)
                        it = min(max(i0[1] + dt_, 0), nt - 1)
                        ip = (i0[2] + dp_) % npz
                        wgt = (fr[0] if dr_ else 1 - fr[0]) * (fr[1] if dt_ else 1 - fr[1]) * (fr[2] if dp_ else 1 - fr[2])
                        acc += wgt * float(cells[ip, it, ir, f])
            out[f] = acc
        return out

    flat = np.ascontiguousarray(cells.reshape(-1))
    for _ in range(300):
        r, u, phi = rng.uniform(1.0, 15.0), rng.uniform(-1, 1), rng.uniform(-20, 20)
        got = np.zeros(4)
        L.orc_volume_fetch(C.byref(vol), flat.ctypes.data_as(C.POINTER(C.c_float)), r, u, phi, P(got))
        np.testing.assert_allclose(got, brute(r, u, phi), rtol=1e-12, atol=1e-13)


def test_flop_constants_match_counting_oracle():
    exe = os.path.join(ROOT, "oracle", "flopcount")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    got = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert got["rhs"] == flops.RHS
    assert got["rk4_step"] == flops.RK4_STEP
    assert got["dp45_attempt"] == flops.DP45_ATTEMPT
    assert got["projection"] == flops.PROJECTION
    assert got["crossing"] == flops.CROSSING
    assert got["disk_hit"] == flops.DISK_HIT
    assert got["volume_sample"] == flops.VOLUME_SAMPLE
    assert got["volume_step_nosample"] == flops.VOLUME_STEP_NOSAMPLE
    assert got["setup"] == flops.SETUP


def test_oracle_against_committed_reference_vectors(L):
    """tests/golden/ref_vectors.json (written by tests/golden/gen_ref_vectors.py
    from the reference's functions compiled in oracle/_ref): R, R' of
    kerrPotentials, critical curve, synchrotron j (no-Boost K2 = 2 Theta^2),
    Planck, and the a = 0 disk g."""
    vec = json.load(open(os.path.join(GOLD, "ref_vectors.json")))
    for v in vec["kerr_potentials"]:
        r, th, a, xi, eta = v["in"]
        out = np.zeros(6)
        L.orc_potentials(a, xi, eta, r, th, P(out))
        assert abs(out[0] - v["out"][0]) <= 1e-12 * max(abs(v["out"][0]), r**4)
        assert abs(out[1] - v["out"][1]) <= 1e-12 * max(abs(v["out"][1]), r**3)
        assert abs((out[4] - v["out"][2]) - xi * xi) <= 1e-10 * max(1.0, xi * xi)
    for v in vec["critical_params"]:
        got = np.zeros(2)
        L.orc_critical_params(v["in"][0], v["in"][1], P(got))
        np.testing.assert_allclose(got, v["out"], rtol=1e-13, atol=1e-12)
    for v in vec["synch_j"]:
        assert abs(L.orc_synch_j(*v["in"], 1.0) - v["out"]) <= 1e-12 * v["out"]
    for v in vec["planck"]:
        assert abs(L.orc_planck(*v["in"]) - v["out"]) <= 1e-13 * max(v["out"], 1e-300)
    for v in vec["disk_g"]:
        r, phi, a, inc = v["in"]
        assert abs(L.orc_disk_g(r, 0.0, r * math.sin(phi) * math.sin(inc)) - v["out"]) <= 1e-13 * v["out"]


# ---------------------------------------------------------------------------
# Disk intensity, LUT sampling, NT spin term, fluid redshift and the volume
# frame transforms
# ---------------------------------------------------------------------------

H_CGS = 2.0 * math.pi * 1.054571817e-27  # Planck's constant, erg s
KB_CGS = 1.380649e-16
C_CGS = 2.99792458e10


@pytest.mark.parametrize("a", [0.5, 0.9, 0.998])
def test_disk_lut_spin_term_matches_reference_diskflux(L, ref, a):
    """thin_disk.h:243-266 diskFlux, Kerr branch: (1 - sqrt(r_in/r))
    (1 + a/2 sqrt(M/r)) / r^3 up to a constant.  The oracle's table, at the
    reference's own range [r_isco, 4 r_isco] and at the c1 disk [r_isco, 20],
    equals the reference's flux at the same radii divided by its largest
    sample.  A dropped spin factor, a sign error in it or sqrt(r) in place of
    sqrt(1/r) each move the normalised table by more than 1e-3."""
    r_in = L.orc_isco(a)
    for r_out in (4.0 * r_in, 20.0):
        n = 256
        got = np.zeros(n)
        assert L.orc_disk_lut(a, r_in, r_out, n, P(got)) == 0
        rr = r_in + (np.arange(n) / (n - 1.0)) * (r_out - r_in)
        theirs = np.array([ref.ref_disk_flux(r, a, r_in, r_out * (1 + 1e-13)) for r in rr])
        theirs /= theirs.max()
        np.testing.assert_allclose(got, theirs, rtol=1e-12, atol=1e-15)
        # the spin factor is visible at this tolerance
        no_spin = (1 - np.sqrt(r_in / rr)) / rr**3
        assert np.max(np.abs(no_spin / no_spin.max() - got)) > 1e-3


def test_disk_intensity_is_redshifted_planck(L, ref):
    """disk_intensity(F, g, nu, T_max) = g^3 B_{nu/g}(T_max F^1/4): the
    reference's planckFunction at the emitted frequency, carried to the
    observer by its grTransformIntensity (rte_integrator.h:221-230, 465-475).
    Rayleigh-Jeans and Wien limits from independent constants; zero for
    F <= 0 or g <= 0."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        F = 10 ** rng.uniform(-4, 0)
        g = rng.uniform(0.05, 1.8)
        nu = 10 ** rng.uniform(15, 18.5)
        T_max = 10 ** rng.uniform(5.5, 7.5)
        T = T_max * F ** 0.25
        want = ref.ref_gr_transform_intensity(ref.ref_planck(nu / g, T), g)
        got = L.orc_disk_intensity(F, g, nu, T_max)
        if want == 0.0:
            assert got == 0.0
            continue
        # exp(x) turns the rounding of x = h nu / (g k T) into a relative
        # error of about x ulp
        x = H_CGS * nu / (g * KB_CGS * T)
        assert abs(got - want) <= 4e-16 * (x + 100.0) * want, (F, g, nu, T_max)
    # Rayleigh-Jeans: x = h nu / (g k T) << 1  ->  I = 2 nu^2 g k T / c^2 (1 - x/2)
    F, g, T_max = 0.3, 0.7, 1e7
    T = T_max * F ** 0.25
    nu = 1e-6 * g * KB_CGS * T / H_CGS
    rj = 2 * nu**2 * g * KB_CGS * T / C_CGS**2
    assert abs(L.orc_disk_intensity(F, g, nu, T_max) / rj - 1.0) <= 2e-6
    # Wien: x >> 1  ->  I = 2 h nu^3 / c^2 exp(-x)
    nu = 40.0 * g * KB_CGS * T / H_CGS
    wien = 2 * H_CGS * nu**3 / C_CGS**2 * math.exp(-40.0)
    assert abs(L.orc_disk_intensity(F, g, nu, T_max) / wien - 1.0) <= 1e-8
    for F, g in ((0.0, 1.0), (-1.0, 1.0), (0.5, 0.0), (0.5, -0.2)):
        assert L.orc_disk_intensity(F, g, 2.4e17, 1e7) == 0.0


def test_lut_linear_nodes_linearity_and_clamp(L):
    """Linear table sampling over [x0, x1]: node values, midpoints as the mean
    of their nodes, exact on linear data, clamped to the end values outside
    the range (NaN reads the first node).  n in place of n - 1, or an index
    off by one, fails the node check."""
    rng = np.random.default_rng(5)
    for n in (2, 3, 17, 256):
        tab = rng.uniform(-3, 5, n)
        x0, x1 = rng.uniform(-5, 2), rng.uniform(3, 30)
        dx = (x1 - x0) / (n - 1)
        for i in range(n):
            got = L.orc_lut_linear(P(tab), n, x0, x1, x0 + i * dx)
            assert abs(got - tab[i]) <= 1e-12 * np.max(np.abs(tab)), (n, i)
        for i in range(n - 1):
            got = L.orc_lut_linear(P(tab), n, x0, x1, x0 + (i + 0.5) * dx)
            assert abs(got - 0.5 * (tab[i] + tab[i + 1])) <= 1e-12 * np.max(np.abs(tab))
        c0, c1 = rng.uniform(-2, 2), rng.uniform(-2, 2)
        lin = c0 + c1 * np.arange(n)
        for x in rng.uniform(x0, x1, 50):
            assert abs(L.orc_lut_linear(P(lin), n, x0, x1, x) - (c0 + c1 * (x - x0) / dx)) <= 1e-12 * (1 + abs(c1) * n)
        for x in (x0 - 1.0, x0 - 1e6, -math.inf):
            assert L.orc_lut_linear(P(tab), n, x0, x1, x) == tab[0]
        for x in (x1, x1 + 0.5, 1e9, math.inf):
            assert abs(L.orc_lut_linear(P(tab), n, x0, x1, x) - tab[-1]) <= 1e-12 * np.max(np.abs(tab))
        assert L.orc_lut_linear(P(tab), n, x0, x1, math.nan) == tab[0]


def _ell_kepler(r, a):
    """Specific angular momentum -u_phi/u_t of the prograde circular
    equatorial orbit (Bardeen, Press & Teukolsky 1972, M = 1)."""
    sr = math.sqrt(r)
    return (r * r - 2 * a * sr + a * a) / (r * sr - 2 * sr + a)


def _g_from_reference_gcon(ref, r, u, a, xi, ell):
    """Redshift of gas with u_phi = -ell u_t seen by a photon (E = 1, L = xi),
    from the reference's inverse Kerr metric (connection.h:127-178):
    u_t = -1/sqrt(-(g^tt - 2 ell g^tphi + ell^2 g^phiphi)), u^mu = g^mu nu u_nu,
    g = 1 / (u^t - xi u^phi).  None where the orbit is superluminal."""
    G = np.zeros(16)
    ref.ref_kerr_gcon(r, math.acos(u), a, 2.0, P(G))
    gtt, gtp, gpp = G[0], G[3], G[15]
    D = gtt - 2 * ell * gtp + ell * ell * gpp
    if D >= 0:
        return None
    ut_lo = -1.0 / math.sqrt(-D)
    up_lo = -ell * ut_lo
    ut_hi = gtt * ut_lo + gtp * up_lo
    up_hi = gtp * ut_lo + gpp * up_lo
    return 1.0 / (ut_hi - xi * up_hi)


def test_fluid_g_kepler_identity_zamo_and_reference_metric(L, ref):
    """fluid_g(r, u, a, xi, ell), the redshift of the torus gas:
      * on the equator with the Keplerian ell it is the thin-disk g
        (disk_g, pinned against kerrDiskGFactor);
      * at ell = 0 (zero angular momentum observers) it is the closed form
        alpha / (1 - omega xi), alpha^2 = Delta Sigma / A, omega = 2 a r / A;
      * everywhere it equals the redshift built from the reference's
        contravariant metric kerrGcon.
    A wrong sign on g_tphi, or sin^2 dropped from g_phiphi, fails all three."""
    rng = np.random.default_rng(17)
    for a in (0.0, 0.5, 0.9, 0.998):
        r_isco = L.orc_isco(a)
        for r in np.linspace(r_isco * 1.001, 30.0, 25):
            ell = _ell_kepler(r, a)
            for xi in rng.uniform(-9, 9, 6):
                want = L.orc_disk_g(r, a, xi)
                got = L.orc_fluid_g(r, 0.0, a, xi, ell)
                assert abs(got - want) <= 1e-12 * want, (a, r, xi)
    for _ in range(400):
        a = rng.uniform(0, 0.998)
        r = rng.uniform(2.2, 40.0)
        u = rng.uniform(-0.95, 0.95)
        xi = rng.uniform(-8, 8)
        s2 = 1 - u * u
        sig = r * r + a * a * u * u
        dlt = r * r - 2 * r + a * a
        A = (r * r + a * a) ** 2 - a * a * dlt * s2
        if dlt <= 0:
            continue
        om = 2 * a * r / A
        alpha = math.sqrt(dlt * sig / A)
        if 1 - om * xi <= 0:
            continue
        got = L.orc_fluid_g(r, u, a, xi, 0.0)
        assert abs(got - alpha / (1 - om * xi)) <= 1e-12 * got
    checked = 0
    for _ in range(2000):
        a = rng.uniform(0, 0.998)
        r = rng.uniform(2.5, 40.0)
        u = rng.uniform(-0.95, 0.95)
        xi = rng.uniform(-8, 8)
        ell = rng.uniform(-2, 6)
        want = _g_from_reference_gcon(ref, r, u, a, xi, ell)
        got = L.orc_fluid_g(r, u, a, xi, ell)
        if want is None or want <= 0:
            assert got == 0.0
            continue
        assert abs(got - want) <= 1e-10 * want, (a, r, u, xi, ell)
        checked += 1
    assert checked > 500


def test_volume_point_frame_transforms_match_reference(L, ref):
    """volume_point's observer-frame coefficients per unit affine parameter:
    e = g^2 j_nu(nu/g) L and a = alpha_nu(nu/g) L / g, i.e. the reference's
    grTransformEmission / grTransformAbsorption (rte_integrator.h:465-515) of
    its synchrotronThermalEmissivity / Absorption at the fluid-frame
    frequency, with g from the reference's inverse metric.  Uniform cells
    (the trilinear fetch is exact on them) with the K_2 ratio 1, the
    reference's no-Boost K_2 = 2 Theta^2; outside the support no sample is
    taken."""
    rng = np.random.default_rng(23)
    n = 4
    L_cm = 1.4766250e5
    for _ in range(200):
        ne, th, B = (np.float32(10 ** rng.uniform(3, 7)), np.float32(10 ** rng.uniform(-0.5, 2)),
                     np.float32(10 ** rng.uniform(0, 2)))
        cells = np.empty((n, n, n, 4), dtype=np.float32)
        cells[...] = (ne, th, B, 1.0)
        flat = np.ascontiguousarray(cells.reshape(-1))
        a = rng.uniform(0, 0.998)
        ell = rng.uniform(0, 4.5)
        vol = abi.BH_Volume(None, n, n, n, 0, 2.0, 50.0, 2.5, 40.0, 0.9, ell)
        r, u, phi = rng.uniform(3, 39), rng.uniform(-0.85, 0.85), rng.uniform(0, 6.28)
        xi, eta = rng.uniform(-6, 6), rng.uniform(0, 30)
        nu = 2.3e11
        ae = np.zeros(2)
        assert L.orc_volume_point(C.byref(vol), flat.ctypes.data_as(C.POINTER(C.c_float)),
                                  a, xi, eta, r, u, phi, nu, L_cm, P(ae)) == 1
        g = _g_from_reference_gcon(ref, r, u, a, xi, ell)
        if g is None or g <= 0:
            assert ae[0] == 0.0 and ae[1] == 0.0
            continue
        nf = nu / g
        e_ref = ref.ref_gr_transform_emission(ref.ref_synch_j(nf, float(B), float(ne), float(th)), g) * L_cm
        a_ref = ref.ref_gr_transform_absorption(ref.ref_synch_alpha(nf, float(B), float(ne), float(th)), g) * L_cm
        assert abs(ae[1] - e_ref) <= 1e-10 * e_ref
        assert abs(ae[0] - a_ref) <= 1e-10 * a_ref
    vol = abi.BH_Volume(None, n, n, n, 0, 2.0, 50.0, 2.5, 40.0, 0.9, 3.0)
    for r, u in ((2.2, 0.0), (45.0, 0.0), (10.0, 0.95)):
        ae = np.ones(2)
        assert L.orc_volume_point(C.byref(vol), flat.ctypes.data_as(C.POINTER(C.c_float)),
                                  0.5, 1.0, 3.0, r, u, 0.0, 2.3e11, L_cm, P(ae)) == 0
        assert ae[0] == 0.0 and ae[1] == 0.0
