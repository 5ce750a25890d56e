"""The Hamiltonian form (kerrGcon, kerrHamiltonian, kerrVelocity, kerrForce,
yoshida4Step, integrateGeodesic), the conserved quantities (computeConserved,
GeodesicConservationMonitor) and the event detection (crossedRadius, hasTurningPoint,
crossedEquator, bisectRadialCrossing, bisectTurningPoint, photonSubringIndex)
of the physics interface (include/bh_physics.hpp) on the CPU.

Pinned by the reference's own functions compiled from
src/physics/symplectic_integrator.h and connection.h into oracle/_ref
(skipped where absent), by H = 0 on null states built from the oracle's
potentials, by finite differences of H (velocity = dH/dp exactly, since H is
quadratic in p; force = -dH/dq to the differences' truncation), and by the
conservation of the cyclic momenta p_t and p_phi.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
_PD = C.POINTER(C.c_double)
_D = C.c_double


def P(a):
    return a.ctypes.data_as(_PD)


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = tmp_path_factory.mktemp("ham") / "physics_api_check"
    lib_dir = os.path.join(ROOT, "blackhole_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "physics_api_check.cpp"), "-o", str(exe),
                    "-L", lib_dir, "-lbhk", f"-Wl,-rpath,{lib_dir}"], check=True)

    def run(what, rows):
        stdin = "\n".join(" ".join(f"{v:.17g}" for v in r) for r in rows)
        out = subprocess.run([str(exe), what], input=stdin, capture_output=True, text=True, check=True).stdout
        rows = [np.array([float(v) for v in l.split()]) for l in out.strip().split("\n")]
        return np.array(rows) if len({len(r) for r in rows}) == 1 else rows
    return run


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libref.so not built (needs /root/reference)")
    R = C.CDLL(REF_LIB)
    R.ref_kerr_gcon.argtypes = [_D, _D, _D, _D, _PD]
    R.ref_kerr_hamiltonian.argtypes = [_PD, _D, _D]
    R.ref_kerr_hamiltonian.restype = _D
    R.ref_kerr_force.argtypes = [_PD, _D, _D, _PD]
    R.ref_yoshida4.argtypes = [_PD, _D, _D, _D, C.c_int]
    return R


def null_state(a, r, th, xi, eta, inbound=True, up=True):
    """Photon (E = 1, L = xi, Carter eta) at (r, theta): covariant momenta from
    the oracle's potentials, p_r = +-sqrt(R) / Delta, p_theta = +-sqrt(Theta)."""
    pot = np.zeros(6)
    OL.lib().orc_potentials(a, xi, eta, r, th, OL.dptr(pot))
    R = pot[0]
    Th = eta + a * a * math.cos(th) ** 2 - xi * xi / math.tan(th) ** 2
    delta = r * r - 2.0 * r + a * a
    pr = (-1.0 if inbound else 1.0) * math.sqrt(R) / delta
    pth = (1.0 if up else -1.0) * math.sqrt(Th)
    return [0.0, r, th, 0.0, -1.0, pr, pth, xi]


def random_states(rng, n):
    rows = []
    for _ in range(n):
        a = rng.uniform(0.0, 0.99)
        st = [rng.uniform(-5, 5), rng.uniform(3.0, 40.0), rng.uniform(0.2, 2.9), rng.uniform(0, 6),
              -rng.uniform(0.5, 2.0), rng.normal(), rng.normal() * 3, rng.normal() * 4]
        rows.append([a, 2.0] + st)
    return rows


def test_null_states_have_zero_hamiltonian(prog):
    rows = []
    for a, r, th, xi, eta in ((0.9, 20.0, 1.0, 3.0, 20.0), (0.5, 8.0, 0.6, -2.0, 12.0), (0.0, 30.0, 1.4, 5.0, 4.0),
                              (0.998, 6.0, 1.2, 1.5, 25.0)):
        rows.append([a, 2.0] + null_state(a, r, th, xi, eta))
    out = prog("hamilton", rows)
    scale = np.abs(out[:, :16]).max(axis=1) * 400.0
    assert np.all(np.abs(out[:, 16]) <= 1e-14 * scale), out[:, 16]


def test_velocity_and_force_are_derivatives_of_h(prog):
    rng = np.random.default_rng(31)
    rows = random_states(rng, 20)
    base = prog("hamilton", rows)
    for row, b in zip(rows, base):
        H0, vel, force = b[16], b[17:21], b[21:25]
        # dH/dp: H is quadratic in p, so the central difference is exact to rounding
        for i in range(4):
            e = 1e-3
            rp, rm = list(row), list(row)
            rp[6 + i] += e
            rm[6 + i] -= e
            hp, hm = prog("hamilton", [rp, rm])[:, 16]
            assert abs((hp - hm) / (2 * e) - vel[i]) <= 1e-9 * max(1.0, abs(vel[i]))
        # -dH/dq for r and theta (fourth-order central difference), 0 for t and phi
        for i, e in ((1, 1e-3), (2, 1e-4)):
            pts = []
            for k in (-2, -1, 1, 2):
                rr = list(row)
                rr[2 + i] += k * e
                pts.append(rr)
            h = prog("hamilton", pts)[:, 16]
            d = (h[0] - 8 * h[1] + 8 * h[2] - h[3]) / (12 * e)
            assert abs(-d - force[i]) <= 1e-7 * max(1.0, abs(force[i])), (i, -d, force[i])
        assert force[0] == 0.0 and force[3] == 0.0


def test_metric_hamiltonian_and_force_match_reference(prog, ref):
    rng = np.random.default_rng(32)
    rows = random_states(rng, 40)
    out = prog("hamilton", rows)
    for row, o in zip(rows, out):
        a, rS, st = row[0], row[1], np.array(row[2:])
        g = np.zeros(16)
        ref.ref_kerr_gcon(st[1], st[2], a, rS, P(g))
        np.testing.assert_allclose(o[:16], g, rtol=1e-14, atol=1e-300)
        h = ref.ref_kerr_hamiltonian(P(st), a, rS)
        assert abs(o[16] - h) <= 1e-13 * max(1.0, np.abs(o[:16]).max() * (st[4:] ** 2).sum())
        f = np.zeros(4)
        ref.ref_kerr_force(P(st), a, rS, P(f))
        # the reference differentiates numerically (eps = 1e-5, central): truncation ~ eps^2
        assert np.all(np.abs(o[21:25] - f) <= 1e-7 * np.maximum(1.0, np.abs(f)))


def test_yoshida_composition_matches_reference_and_is_first_order(prog, ref):
    a, rS = 0.9, 2.0
    st = null_state(a, 20.0, 1.0, 3.0, 20.0)
    res = {}
    for h in (0.4, 0.2, 0.1):
        n = int(round(8.0 / h))
        mine = prog("yoshida", [[a, rS, h, n] + st])[0]
        theirs = np.array(st, dtype=np.float64)
        ref.ref_yoshida4(P(theirs), a, rS, h, n)
        # same composition; only the force differs (analytic vs eps = 1e-5 differences)
        np.testing.assert_allclose(mine[:8], theirs, rtol=1e-7, atol=1e-9)
        assert mine[4] == st[4] and mine[7] == st[7]     # p_t, p_phi: cyclic, exactly conserved
        res[h] = mine
    dH = [abs(res[h][8]) for h in (0.4, 0.2, 0.1)]
    dr = [abs(res[0.4][1] - res[0.2][1]), abs(res[0.2][1] - res[0.1][1])]
    assert 1.8 < dH[0] / dH[1] < 2.2 and 1.8 < dH[1] / dH[2] < 2.2     # first order in h
    assert 1.8 < dr[0] / dr[1] < 2.2


def test_event_detection(prog):
    """Bisection on analytic rays finds the crossing of r = 8 by
    r = 12 - 0.7 l - 0.01 l^3 and the turning point of r = 10 + (l - 2.3)^2
    to their tolerances; the predicates and the subring index."""
    rows = prog("events", [[0]])
    e1, e2, e3 = rows[0], rows[1], rows[2]
    # l* solves 0.01 l^3 + 0.7 l - 4 = 0
    ls = np.roots([0.01, 0.0, 0.7, -4.0])
    ls = float(ls[np.abs(ls.imag) < 1e-12].real[0])
    assert e1[0] == 4 and e1[5] == 1 and abs(e1[2] - 8.0) < 1e-10 and abs(e1[1] - ls) < 2e-10
    assert abs(e1[3] - (1.0 + 0.1 * e1[1])) < 1e-15
    assert e2[0] == 3 and e2[5] == 1 and abs(e2[1] - 2.3) <= 1e-12
    assert e3[4] == 7 and abs(e3[1] - ls) <= 10.0 / 2 ** 8        # budget: 7 halvings, then the midpoint
    assert list(rows[3]) == [1, 0, 1, 0, 1, 0, 3]


def test_event_detection_matches_reference(prog, ref):
    ref.ref_events.argtypes = [_PD]
    want = np.zeros(25)
    ref.ref_events(P(want))
    rows = prog("events", [[0]])
    got = np.concatenate([rows[0], rows[1], rows[2], rows[3]])
    np.testing.assert_array_equal(got, want)


def test_reference_symplectic_known_answers(prog):
    """tests/symplectic_integrator_test.cpp: the timelike orbit from r = 10 M
    (a = 0.5, p = (-1, 0, 0, 4.878)) has H near -1/2; 1000 steps of h = 0.01
    keep |H_N - H_0| < 1e-6 and the cyclic momenta to 1e-10; 500 steps
    forward then 500 back return to the start within 1e-4."""
    st = [0.0, 10.0, math.pi / 2, 0.0, -1.0, 0.0, 0.0, 4.878]
    a, rS = 0.5, 2.0
    H0 = prog("hamilton", [[a, rS] + st])[0][16]
    assert abs(H0 + 0.5) <= 5e-3 and H0 < 0.0
    out = prog("integrate", [[a, rS, 10.0, 1000] + st])[0]
    assert abs(out[8] - H0) < 1e-6, out[8] - H0
    assert abs(out[4] - st[4]) < 1e-10 and abs(out[7] - st[7]) < 1e-10
    fwd = prog("integrate", [[a, rS, 5.0, 500] + st])[0]
    back = prog("integrate", [[a, rS, -5.0, 500] + list(fwd[:8])])[0]
    assert np.abs(back[:8] - np.array(st)).max() < 1e-4
    assert abs(back[4] - st[4]) < 1e-14 and abs(back[7] - st[7]) < 1e-14


def _null_velocity(a, r, th, xi, eta, sr=-1.0, sth=1.0):
    """Contravariant velocity u^mu = g^{mu nu} p_nu of the photon (E = 1,
    L = xi, Carter eta) at (r, theta)."""
    st = null_state(a, r, th, xi, eta, inbound=sr < 0, up=sth > 0)
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    sig, dl = r * r + a * a * c2, r * r - 2 * r + a * a
    A = (r * r + a * a) ** 2 - a * a * dl * s2
    pt, pr, pth, pph = st[4:]
    ut = -A / (dl * sig) * pt - 2 * r * a / (dl * sig) * pph
    uph = -2 * r * a / (dl * sig) * pt + (dl - a * a * s2) / (dl * sig * s2) * pph
    return [r, th, ut, dl / sig * pr, pth / sig, uph]


def test_conserved_quantities_round_trip_and_monitor(prog):
    """A photon built from (E, L, Q) = (1, xi, eta) gives them back, with
    mu2 = 0; the monitor's drift bookkeeping."""
    cases = ((0.9, 20.0, 1.0, 3.0, 20.0), (0.5, 8.0, 0.6, -2.0, 12.0), (0.0, 30.0, 1.4, 5.0, 4.0))
    rows = [_null_velocity(a, r, th, xi, eta) + [1.0, a] for a, r, th, xi, eta in cases]
    out = prog("conserved", rows)
    for (a, r, th, xi, eta), o in zip(cases, out[:3]):
        assert abs(o[0] - 1.0) <= 1e-12 and abs(o[1] - xi) <= 1e-12 * max(1, abs(xi))
        assert abs(o[2] - eta) <= 1e-10 * eta and abs(o[3]) <= 1e-12
    # the monitor: first row initial, later rows drift relative to it
    base = _null_velocity(0.9, 20.0, 1.0, 3.0, 20.0) + [1.0, 0.9]
    moved = list(base)
    moved[2] *= 1.0 + 1e-6     # u^t off by 1e-6: E, L, mu2 drift
    out = prog("conserved", [base, base, moved])
    last, dE, dL, dQ, dM, steps, viol = out[-1]
    assert steps == 2 and viol == 1 and dE > 1e-7 and last == max(dE, dL, dQ, dM)


def test_conserved_quantities_match_reference(prog, ref):
    ref.ref_compute_conserved.argtypes = [_D] * 8 + [_PD]
    rng = np.random.default_rng(33)
    rows = []
    for _ in range(30):
        rows.append([rng.uniform(3, 50), rng.uniform(0.2, 2.9), rng.uniform(0.5, 2.0), rng.normal(),
                     rng.normal() * 0.1, rng.normal() * 0.05, 1.0, rng.uniform(0, 0.99)])
    out = prog("conserved", rows)[:-1]
    for row, o in zip(rows, out):
        want = np.zeros(4)
        ref.ref_compute_conserved(*row, P(want))
        np.testing.assert_allclose(o, want, rtol=1e-12, atol=1e-12)
