"""The two measured references of the roofline: the streaming copy kernel
(phz_copy_kernel) moves its bytes and reaches a plausible share of the 8 TB/s
HBM3E peak, and the MFMA microbenchmark (phz_mfma_peak) reports rates below
MI355X's published dense peaks (f16 / bf16 2.5 PFLOP/s, f32 157.3 TFLOP/s)
and well above what a single CU could issue."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_copy_kernel_copies_and_reaches_hbm_rates():
    from phaser_amd.capi import Context
    with Context(bandwidth=8, n_voxels=8, max_points=16) as c:
        n = 1 << 28
        src = np.arange(n // 4, dtype=np.uint32)
        a = c.device_array(src.view(np.float32))
        b = c.device_empty(n)
        ms = c.copy_kernel(b.ptr, a.ptr, n, 5)
        back = b.download((n // 4,), np.uint32)
        a.free()
        b.free()
    gbs = 2 * n / (ms * 1e-3) / 1e9
    print(f"copy kernel {gbs:.0f} GB/s", flush=True)
    assert np.array_equal(back, src)
    assert 1000 < gbs < 8000


@pytest.mark.parametrize("dtype,peak", [(0, 2500.0), (1, 2500.0), (2, 157.3)])
def test_mfma_peak_is_plausible(dtype, peak):
    from phaser_amd.capi import Context
    with Context(bandwidth=8, n_voxels=8, max_points=16) as c:
        t = max(c.mfma_peak(dtype, 4000) for _ in range(3))
    print(f"dtype {dtype}: {t:.1f} TFLOP/s of {peak}", flush=True)
    assert 0.3 * peak < t < 1.05 * peak
