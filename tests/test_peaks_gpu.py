"""rbc_measure_peaks: the measured BF16 MFMA and FP32 FMA rates of the device fall inside
what the MI355X can deliver (2.5 PFLOP/s dense BF16, 157.3 TFLOP/s FP32 vector) and well
above what a loop with extra instructions in it would show."""
import pytest

from conftest import rbc

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible():
    p = rbc.measure_peaks()
    assert 1500.0 < p["bf16_mfma_tflops"] < 2600.0, p
    assert 60.0 < p["fp32_fma_tflops"] < 165.0, p
