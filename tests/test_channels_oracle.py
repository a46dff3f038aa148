"""The oracle's reflectivity and ambient channels (clouds read from PLY
files carry them, point-cloud.cc:152-155 and phase-aligner.cc:113-118), on
the CPU, pinned by invariants of the sampling and voxel maps: a channel
equal to the intensity (or the range) gives exactly the intensity's (the
range's) grid, zero channels equal no channels, and live channels change
the correlation (the four-way fusions are no longer degenerate)."""
import numpy as np

from oracle import oracle
from phaser_amd import synth


def extras(cloud):
    """Scene-consistent channels for tests: functions of the intensity."""
    i = cloud[:, 3].astype(np.float32)
    return np.stack([i * i / np.float32(100), np.float32(50) + np.float32(0.5) * i], 1)


def _small_pair(seed=0):
    src, tgt, _, _ = synth.pair(seed=seed, rings=32, azimuths=256)
    return src, tgt


def test_sampling_extra_channels():
    src, _ = _small_pair()
    base = oracle.sample_cloud(16, src)
    np.testing.assert_array_equal(oracle.sample_cloud(16, src, extra=np.zeros((len(src), 2))), base)
    same = np.stack([src[:, 3], src[:, 3]], 1)
    g = oracle.sample_cloud(16, src, extra=same)
    np.testing.assert_array_equal(g[:2], base[:2])
    np.testing.assert_array_equal(g[2], base[1])
    np.testing.assert_array_equal(g[3], base[1])


def test_voxel_extra_channels():
    src, _ = _small_pair(1)
    p = src[:, :3].astype(np.float32)
    rng32 = np.sqrt(p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1] + p[:, 2] * p[:, 2])  # float32 sqrtf
    base = oracle.voxelize(src, 32)
    ch = oracle.voxelize(src, 32, extra=np.stack([src[:, 3], rng32], 1))
    np.testing.assert_array_equal(ch[:2], base[:2])
    np.testing.assert_array_equal(ch[2], base[0])
    np.testing.assert_array_equal(ch[3], base[1])
    assert not base[2].any() and not base[3].any()


def test_register_with_channels():
    src, tgt = _small_pair(2)
    base = oracle.register(tgt, src, bw=16, n_voxels=32, want_grids=True)
    zero = oracle.register(tgt, src, bw=16, n_voxels=32, want_grids=True,
                           target_extra=np.zeros((len(tgt), 2)),
                           source_extra=np.zeros((len(src), 2)))
    np.testing.assert_array_equal(zero["rot_corr"], base["rot_corr"])
    np.testing.assert_array_equal(zero["trans_corr"], base["trans_corr"])
    live = oracle.register(tgt, src, bw=16, n_voxels=32, want_grids=True,
                           target_extra=extras(tgt), source_extra=extras(src))
    assert np.max(np.abs(live["rot_corr"] - base["rot_corr"])) > 1e-3 * np.max(base["rot_corr"])
    assert np.max(np.abs(live["trans_corr"] - base["trans_corr"])) > \
        1e-3 * np.max(np.abs(base["trans_corr"]))
