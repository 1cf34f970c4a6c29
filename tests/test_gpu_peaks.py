"""Measured device peaks behind the rooflines (pt_peak_probe): a streaming
copy kernel and an FP32 fused multiply-add loop, timed with HIP events."""
import pytest

pytestmark = pytest.mark.gpu


def test_peak_probe_rates_are_plausible(gpu_ctx):
    """A streaming copy between 1 and 8 TB/s (MI355X HBM3E: 8 TB/s published)
    and the unpacked FP32 FMA loop between 10 and 157.3 TFLOP/s (the published
    FP32 vector peak, which counts packed FP32)."""
    bw, nbytes, ms = gpu_ctx.peak_probe(0, 1 << 26, 5)
    assert nbytes == 8 * (1 << 26) and ms > 0
    assert 1e12 < bw < 8e12, bw
    fl, nflop, ms = gpu_ctx.peak_probe(1, 4096, 5)
    # 16 flops per round and thread, 16 blocks of 256 threads per CU
    assert nflop > 0 and nflop % (16 * 4096 * 256 * 16) == 0 and ms > 0
    assert 1e13 < fl < 157.3e12, fl
