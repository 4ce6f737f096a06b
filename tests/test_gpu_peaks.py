"""GPU test of the attainable-peak diagnostic (helix_hip_measure_peaks): every figure is positive and no figure
exceeds the MI355X datasheet peak it is measured against (8 TB/s HBM for reads and for a copy counted as bytes read
plus bytes written, 78.6 / 157.3 TFLOP/s dense FP64 / FP32 matrix) by more than the clock's boost margin; bad
arguments are contract errors."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_peaks_are_positive_and_below_the_datasheet():
    import helix_amd as hx
    pk = hx.measure_peaks(0, 1 << 30, 3)
    print(pk)
    assert pk["bytes"] == 1 << 30 and pk["compute_units"] > 0
    for key, datasheet in (("hbm_read_gbs", 8000.0), ("hbm_copy_gbs", 8000.0), ("mfma_f64_tflops", 78.6),
                           ("mfma_f32_tflops", 157.3)):
        assert 0.05 * datasheet < pk[key] <= 1.1 * datasheet, (key, pk[key])


def test_peaks_reject_bad_arguments():
    from helix_amd import capi
    out = capi.Peaks()
    lib = capi.hip()
    assert lib.helix_hip_measure_peaks(0, 1 << 19, 3, C.byref(out)) == -1
    assert b"1 MiB <= bytes" in lib.helix_hip_last_error()
    assert lib.helix_hip_measure_peaks(0, 1 << 20, 0, C.byref(out)) == -1
    assert lib.helix_hip_measure_peaks(99, 1 << 20, 3, C.byref(out)) == -1
    assert b"no such device" in lib.helix_hip_last_error()
