"""smc_calibrate_fma: the plain-FMA rate of the device that the VALU
kernels' rooflines are read against.  One wave64 FP32 FMA per SIMD every
4 cycles on 1024 SIMDs is 78.6 TFLOP/s at 2.4 GHz; a chain of independent
FMAs lands below that and well above a tenth of it."""
import pytest

pytestmark = pytest.mark.gpu


def test_calibrate_fma_reads_the_plain_fma_rate():
    from supermc_amd import _native
    tf, ms = _native.calibrate_fma()
    assert ms > 0.0
    assert 10.0 < tf < 80.0, tf
    print(f"plain FP32 FMA chain: {tf:.1f} TFLOP/s in {ms:.3f} ms", flush=True)
