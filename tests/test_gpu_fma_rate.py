"""The device FMA-rate probe behind the measured peak in bench.py: it must
report a rate of the order of MI355X's published dense FP64 vector peak
(78.6 TFLOP/s), and the unpacked FP32 form (all device code here is built
without packed FP32) at most twice that."""
import pytest

from octotiger_amd import _abi, native

pytestmark = pytest.mark.gpu

PUBLISHED_FP64_TFLOPS = 78.6


def test_fma_rate_probe_reports_device_peaks():
    ctx = native.Context(theta=0.5, device=0)
    fp64 = ctx.measure_fma_rate(_abi.PRECISION_FP64)
    fp32 = ctx.measure_fma_rate(_abi.PRECISION_MIXED)
    ctx.close()
    print(f"dense FMA rate: fp64 {fp64:.1f} TFLOP/s, fp32 (unpacked) {fp32:.1f} TFLOP/s", flush=True)
    assert 0.3 * PUBLISHED_FP64_TFLOPS < fp64 < 1.1 * PUBLISHED_FP64_TFLOPS
    assert 0.3 * PUBLISHED_FP64_TFLOPS < fp32 < 2.2 * PUBLISHED_FP64_TFLOPS


def test_fma_rate_probe_rejects_bad_precision():
    ctx = native.Context(theta=0.5, device=0)
    with pytest.raises(native.OctofmmError):
        ctx.measure_fma_rate(7)
    ctx.close()
