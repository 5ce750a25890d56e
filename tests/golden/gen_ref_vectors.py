#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.json: input/output vectors of the
reference's own FP64 functions, evaluated by oracle/_ref/libref.so (built by
oracle/build_ref.sh from /root/reference/src/physics).  Run in the container
that holds the reference; the JSON is the committed fixture."""
import ctypes as C
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libref.so")
_D, _PD = C.c_double, C.POINTER(C.c_double)


def main():
    R = C.CDLL(LIB)
    R.ref_kerr_potentials.argtypes = [_D] * 6 + [_PD]
    R.ref_critical_params.argtypes = [_D, _D, _PD]
    R.ref_photon_orbit.argtypes = [_D, C.c_int]
    R.ref_photon_orbit.restype = _D
    R.ref_synch_j.argtypes = [_D] * 4
    R.ref_synch_j.restype = _D
    R.ref_planck.argtypes = [_D, _D]
    R.ref_planck.restype = _D
    R.ref_disk_g.argtypes = [_D] * 4
    R.ref_disk_g.restype = _D
    rng = np.random.default_rng(20261203)
    out = {"kerr_potentials": [], "critical_params": [], "synch_j": [], "planck": [], "disk_g": []}
    for _ in range(64):
        a, xi, eta = rng.uniform(0, 0.999), rng.uniform(-8, 8), rng.uniform(0, 40)
        r, th = rng.uniform(1.5, 80), rng.uniform(0.2, 2.9)
        o = np.zeros(4)
        R.ref_kerr_potentials(r, th, a, 1.0, xi, eta, o.ctypes.data_as(_PD))
        out["kerr_potentials"].append({"in": [r, th, a, xi, eta], "out": o.tolist()})
    for a in (0.3, 0.5, 0.9, 0.94, 0.998):
        rp, rm = R.ref_photon_orbit(a, 1), R.ref_photon_orbit(a, 0)
        for r in np.linspace(rp + 1e-3, rm - 1e-3, 7):
            o = np.zeros(2)
            R.ref_critical_params(r, a, o.ctypes.data_as(_PD))
            out["critical_params"].append({"in": [float(r), a], "out": o.tolist()})
    for _ in range(32):
        nu, B = 10 ** rng.uniform(9, 13), 10 ** rng.uniform(-1, 2)
        ne, th = 10 ** rng.uniform(-2, 7), 10 ** rng.uniform(-0.5, 2.5)
        out["synch_j"].append({"in": [nu, B, ne, th], "out": R.ref_synch_j(nu, B, ne, th)})
        T = 10 ** rng.uniform(3, 12)
        out["planck"].append({"in": [nu, T], "out": R.ref_planck(nu, T)})
    for r in (6.5, 8.0, 15.0):
        for phi in (0.3, 2.8):
            for inc in (0.2, 1.5):
                out["disk_g"].append({"in": [r, phi, 0.0, inc], "out": R.ref_disk_g(r, phi, 0.0, inc)})
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
