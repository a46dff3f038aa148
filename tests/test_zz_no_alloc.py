"""phz_spatial_correlate_single_noalloc (no device allocation, a file-scope
device array) against phz_spatial_correlate_single on the same inputs,
through the C ABI: two calls in a row, the second at a smaller size, then
the argument checks of the no-allocation path.

Both entry points run the same kernels on the same twiddle tables, so the
grids and argmaxes must be identical (bit for bit); the grids are also held
to the oracle's FP64 correlation (spatial-correlation.cc:163-206) within
1e-5 of the grid's maximum, the single-channel parity bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _signals(N, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros(N ** 3, np.float32)
    g = np.zeros(N ** 3, np.float32)
    # sparse voxel means, as a voxelised scan gives them
    idx = rng.choice(N ** 3, size=N ** 3 // 10, replace=False)
    f[idx] = rng.uniform(0.0, 1.0, idx.size)
    shift = rng.integers(0, N, 3)
    g = np.roll(f.reshape(N, N, N), tuple(shift), axis=(0, 1, 2)).reshape(-1).copy()
    g += rng.normal(0.0, 0.05, g.size).astype(np.float32) * (g != 0)
    return f, g


def test_noalloc_matches_the_allocating_entry_point(po):
    from phaser_amd.capi import Context, PhaserError
    with Context(bandwidth=16, n_voxels=64, max_points=16) as ctx:
        for N, seed in ((96, 0), (48, 1)):  # second call smaller: stale contents stay behind
            print(f"no-alloc case N={N} seed={seed}", flush=True)
            f, g = _signals(N, seed)
            ref_corr, ref_am = ctx.spatial_correlate_single(f, g, N)
            d_f, d_g = ctx.device_array(f), ctx.device_array(g)
            d_corr = ctx.device_empty(4 * N ** 3)
            am = ctx.spatial_correlate_single_noalloc(d_f, d_g, d_corr, N)
            corr = d_corr.download(N ** 3, np.float32)
            assert am == ref_am
            assert np.array_equal(corr, ref_corr)
            # without a grid output: the same argmax
            assert ctx.spatial_correlate_single_noalloc(d_f, d_g, None, N) == ref_am
            o = po.spatial_correlate(f.astype(np.float64), g.astype(np.float64), N)
            assert np.max(np.abs(corr - o)) <= 1e-5 * np.max(np.abs(o))
            assert am == int(np.argmax(o))
            for d in (d_f, d_g, d_corr):
                d.free()
        # sizes beyond the array's maximum are refused before any launch
        d = ctx.device_empty(4 * 8)
        for N in (129, 256, 1):
            with pytest.raises(PhaserError):
                ctx.spatial_correlate_single_noalloc(d, d, None, N)
        d.free()
