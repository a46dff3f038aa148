"""Spherical zero padding in the oracle (the SO(3) correlation on the grid
of bw + padding from coefficients of degree < bw, spherical-correlation.cc:
21-27 and 248-278), on the CPU.  Pinned end to end like the unpadded
chain: a cloud and its exact rotation register to that rotation within the
finer bin of the padded grid (the grid layout, index decode and pose all
at bw + padding), zero translation; padding 0 is the unpadded result."""
import math

import numpy as np

from oracle import oracle
from phaser_amd import synth


def _rot_angle(Ra, Rb):
    c = (np.trace(Ra.T @ Rb) - 1) / 2
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def test_zero_padding_is_the_unpadded_result():
    s, t, _, _ = synth.pair(seed=2, rings=32, azimuths=256)
    a = oracle.register(t, s, bw=16, n_voxels=32, want_grids=True)
    b = oracle.register(t, s, bw=16, n_voxels=32, want_grids=True, sph_pad=0)
    np.testing.assert_array_equal(a["rot_corr"], b["rot_corr"])


def test_padded_grid_recovers_a_pure_rotation():
    """bw 32 padded to 48: bins of 3.75 deg instead of 5.6 deg."""
    s, _, R, _ = synth.pair(seed=0, **synth.CONFIG_B32)
    tg = s.copy()
    tg[:, :3] = (s[:, :3].astype(np.float64) @ R.T).astype(np.float32)
    r = oracle.register(tg, s, bw=32, n_voxels=64, sph_pad=16, want_grids=True)
    assert r["rot_corr"].shape == (96 ** 3,)
    T = r["T_rot"].astype(np.float64)[:3, :3]
    err = _rot_angle(T, R)
    print(f"padded rotation error {err:.2f} deg")
    assert err < 3.75
    assert np.all(r["translation"] == 0)
    np.testing.assert_allclose(r["zyz"], oracle.zyz_from_index(r["rot_index"], 48))
