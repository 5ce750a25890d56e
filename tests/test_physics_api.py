"""The physics C++ interface (include/bh_physics.hpp) on the CPU: closed
forms and potentials against the oracle, kerrStepMino's order and its
passage through a radial turning point (r = 4.972845 for a = 0.9, b = 6,
the largest root of R), the verified-layer integrators, and the mapping of a
(state, constants) pair to the tracer's image-plane ray."""
from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = tmp_path_factory.mktemp("phys") / "physics_api_check"
    lib_dir = os.path.join(ROOT, "blackhole_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "physics_api_check.cpp"), "-o", str(exe),
                    "-L", lib_dir, "-lbhk", f"-Wl,-rpath,{lib_dir}"], check=True)

    def run(what, stdin=""):
        out = subprocess.run([str(exe), what], input=stdin, capture_output=True, text=True, check=True).stdout
        return np.array([[float(v) for v in l.split()] for l in out.strip().split("\n")])
    return run


def test_closed_forms_match_oracle(prog):
    L = OL.lib()
    for s, rp, ipro, iret, ppro, pret in prog("closed"):
        assert abs(rp - L.orc_outer_horizon(s)) <= 1e-14 * rp
        assert abs(ipro - L.orc_isco(s)) <= 1e-13 * ipro
        assert abs(ppro - L.orc_photon_orbit(s, 1)) <= 1e-13 * ppro
        assert abs(pret - L.orc_photon_orbit(s, 0)) <= 1e-13 * pret
        assert iret > 6.0 - 1e-12 and pret >= 3.0 - 1e-12


def test_potentials_match_oracle(prog):
    rng = np.random.default_rng(4)
    rows = np.stack([rng.uniform(0, 0.99, 50), rng.uniform(-8, 8, 50), rng.uniform(0, 40, 50),
                     rng.uniform(1.5, 50, 50), rng.uniform(0.1, 3.0, 50)], 1)
    got = prog("potentials", "\n".join(" ".join(f"{v:.17g}" for v in r) for r in rows))
    for (a, xi, eta, r, th), (R, dR, T, dT) in zip(rows, got):
        ref = np.zeros(6)
        OL.lib().orc_potentials(a, xi, eta, r, th, OL.dptr(ref))
        assert abs(R - ref[0]) <= 1e-10 * max(1.0, abs(ref[0]))
        assert abs(dR - ref[1]) <= 1e-10 * max(1.0, abs(ref[1]))
        # the reference's thetaPot: the polar potential minus lz^2
        assert abs(T - (ref[4] - xi * xi)) <= 1e-10 * max(1.0, abs(ref[4]))
        assert abs(dT - ref[5]) <= 1e-9 * max(1.0, abs(ref[5]))


def test_step_mino_passes_turning_point_and_is_fourth_order(prog):
    (rmin, turned), = prog("turning")
    assert turned == 1
    assert abs(rmin - 4.972845) <= 2e-6
    (e1, e2), = prog("order")
    assert 12.0 <= e1 / e2 <= 20.0  # RK4: the difference shrinks ~16x per halving


def test_verified_integrators(prog):
    (e_rk4, e_dp, err1, err2, grow), = prog("verified")
    # on y' = y one RK4 step is the degree-4 Taylor polynomial of e^h: its
    # error is exactly the remainder
    h = 0.1
    remainder = math.exp(h) - (1 + h + h * h / 2 + h ** 3 / 6 + h ** 4 / 24)
    assert abs(e_rk4 - remainder) <= 1e-15
    # DP45's fifth-order solution: error below the degree-5 remainder at h = 0.2
    h = 0.2
    assert e_dp <= math.exp(h) - sum(h ** k / math.factorial(k) for k in range(6))
    assert 20.0 <= err1 / err2 <= 45.0  # local error estimate ~h^5
    assert grow == pytest.approx(0.9)


def test_ray_spec_mapping_round_trips_the_image_plane(prog):
    rng = np.random.default_rng(9)
    rows = np.stack([rng.uniform(0, 0.99, 20), rng.uniform(-15, 15, 20), rng.uniform(-15, 15, 20),
                     rng.uniform(0.2, 2.9, 20), rng.uniform(20, 500, 20)], 1)
    got = prog("rayspec", "\n".join(" ".join(f"{v:.17g}" for v in r) for r in rows))
    for (a, al, be, th, r), (ok, al2, be2, sp, th2, r2) in zip(rows, got):
        assert ok == 1
        assert abs(al2 - al) <= 1e-9 * (1 + abs(al)) and abs(be2 - be) <= 1e-7 * (1 + abs(be))
        assert (sp, th2, r2) == (a, th, r)
