"""bh_symplectic.h argument handling on the host (no GPU needed: every case
returns before any device call)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from blackhole_amd._lib import lib


def test_symplectic_rejects_bad_arguments_and_accepts_empty_batches():
    L = lib()
    buf = np.zeros(8)
    p = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert L.bh_symplectic_integrate(None, -1, 1.0, 10, 0.9, 2.0, None) == -1
    assert b"negative" in L.bh_last_error()
    assert L.bh_symplectic_integrate(None, 4, 1.0, 10, 0.9, 2.0, None) == -1
    assert b"NULL" in L.bh_last_error()
    assert L.bh_symplectic_integrate_host(p, 1, 1.0, 10, 0.9, 0.0) == -1
    assert b"rS" in L.bh_last_error()
    assert L.bh_symplectic_integrate_host(p, 1, float("nan"), 10, 0.9, 2.0) == -1
    assert L.bh_symplectic_hamiltonian(None, 3, 0.9, 2.0, None, None) == -1
    # empty batches and non-positive step counts are no-ops
    assert L.bh_symplectic_integrate(None, 0, 1.0, 10, 0.9, 2.0, None) == 0
    assert L.bh_symplectic_integrate_host(p, 1, 1.0, 0, 0.9, 2.0) == 0
    assert L.bh_symplectic_hamiltonian(None, 0, 0.9, 2.0, None, None) == 0
    assert np.all(buf == 0.0)
