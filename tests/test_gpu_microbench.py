"""The measured roofline ceilings (fhe_measure_peaks): positive, and within
the range an MI355X can reach (FP64 vector FMA at most the published 78.6
TFLOP/s plus clock headroom; streaming copy below the 8 TB/s HBM peak)."""

import ctypes

import pytest

import cufhedb_amd as F

pytestmark = pytest.mark.gpu


def test_measured_peaks_plausible(ctx636):
    fp64, copy = ctypes.c_double(), ctypes.c_double()
    F.check(F.lib().fhe_measure_peaks(ctx636.h, ctypes.byref(fp64), ctypes.byref(copy)))
    print(f"\n  FP64 FMA {fp64.value:.1f} TFLOP/s, copy {copy.value:.0f} GB/s")
    assert 5.0 < fp64.value < 90.0
    assert 500.0 < copy.value < 8000.0
