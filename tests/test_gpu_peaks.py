"""bh_measure_peaks through the C ABI: every measured rate positive and
within what MI355X's published peaks allow (FP64 vector 78.6 TFLOP/s,
scalar FP32 the same rate without packed instructions, BF16 dense MFMA
2.5 PFLOP/s, HBM3E 8 TB/s), and at least a third of it (the micro-kernels
keep every SIMD busy; a rate far below would mean the kernel or its FLOP
count is wrong).  NULL output is an error."""
from __future__ import annotations

import ctypes as C

import pytest

from blackhole_amd._lib import lib

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible():
    out = (C.c_double * 4)()
    print("bh_measure_peaks", flush=True)
    assert lib().bh_measure_peaks(out, None) == 0, lib().bh_last_error()
    fp64, fp32, bf16, hbm = list(out)
    print("fp64", fp64, "fp32 scalar", fp32, "bf16 mfma", bf16, "copy GB/s", hbm, flush=True)
    for got, peak in ((fp64, 78.6), (fp32, 78.6), (bf16, 2500.0), (hbm, 8000.0)):
        assert peak / 3.0 <= got <= 1.05 * peak, (got, peak)
    assert lib().bh_measure_peaks(None, None) != 0
