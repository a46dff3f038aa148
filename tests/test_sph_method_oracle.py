"""SphRegistration (CloudController("sph"), sph-registration.cc:82-158) in
the oracle, on the CPU: one sampled channel 0.6 intensity + 0.4 range
(spherical-correlation.cc:54-80, 103-112), transformed and correlated
without the Laplace fusion's shifts, and the refine_rot_x/y/z masks.

Pinned by what the mathematics fixes: a constant function has only the
degree-0 coefficient, so the correlation grid of two constant signals is
constant, C = c f00 h00, and |C|^2 scales with the fourth power of the
blended constant of a self-correlation (fixes both weights and which channel
is intensity: the swapped blend gives the inverse ratio); a signal carried
by intensity alone and by range alone correlate to |C|^2 grids in the ratio
(0.6/0.4)^4 everywhere; a cloud and its
exact rotation register to that rotation within a bin; masking an axis
zeroes that Euler angle of the applied transform only."""
import math

import numpy as np

from oracle import oracle
from phaser_amd import synth


def _grids(rng_val, int_val, n):
    g = np.zeros((4, n * n))
    g[0], g[1] = rng_val, int_val
    return g


def test_constant_signals_scale_with_the_blend():
    bw = 8
    n = 2 * bw
    a = oracle.spherical_correlate_blend(bw, _grids(10.0, 50.0, n), _grids(10.0, 50.0, n))
    b = oracle.spherical_correlate_blend(bw, _grids(50.0, 10.0, n), _grids(50.0, 10.0, n))
    assert np.ptp(a) <= 1e-9 * a.max() and np.ptp(b) <= 1e-9 * b.max()
    # 0.6 * 50 + 0.4 * 10 = 34 against 0.6 * 10 + 0.4 * 50 = 26
    assert math.isclose(a[0] / b[0], (34.0 / 26.0) ** 4, rel_tol=1e-12)


def test_intensity_and_range_weights():
    bw = 8
    n = 2 * bw
    rng = np.random.default_rng(0)
    f = rng.random(n * n) * 20
    h = rng.random(n * n) * 20
    zero = np.zeros(n * n)
    on_i = oracle.spherical_correlate_blend(bw, np.stack([zero, f, zero, zero]),
                                            np.stack([zero, h, zero, zero]))
    on_r = oracle.spherical_correlate_blend(bw, np.stack([f, zero, zero, zero]),
                                            np.stack([h, zero, zero, zero]))
    both = oracle.spherical_correlate_blend(bw, np.stack([f, f, zero, zero]),
                                            np.stack([h, h, zero, zero]))
    np.testing.assert_allclose(on_i, 2.25 ** 2 * on_r, rtol=1e-10)
    # reflectivity and ambient take no part
    np.testing.assert_array_equal(
        oracle.spherical_correlate_blend(bw, np.stack([f, f, h, f]), np.stack([h, h, f, h])), both)
    # the blend of a signal with itself is the signal: C is 1/0.36 times the
    # intensity-only one
    np.testing.assert_allclose(both, on_i / 0.36 ** 2, rtol=1e-10)


def _rot_angle(Ra, Rb):
    c = (np.trace(Ra.T @ Rb) - 1) / 2
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def test_sph_method_recovers_a_pure_rotation_and_masks_axes():
    s, _, R, _ = synth.pair(seed=0, **synth.CONFIG_B32)
    tg = s.copy()
    tg[:, :3] = (s[:, :3].astype(np.float64) @ R.T).astype(np.float32)
    r = oracle.register(tg, s, bw=32, n_voxels=64, method=1, want_grids=True)
    assert r["rot_corr"].shape == (64 ** 3,)
    T = r["T_rot"].astype(np.float64)[:3, :3]
    err = _rot_angle(T, R)
    print(f"sph method rotation error {err:.2f} deg")
    assert err < 5.625
    np.testing.assert_allclose(r["zyz"], oracle.zyz_from_index(r["rot_index"], 32))
    # refine_rot_y false: T = Tx Tz of the same Euler angles
    m = oracle.register(tg, s, bw=32, n_voxels=64, method=1, skip_rot_axes=2)
    assert m["rot_index"] == r["rot_index"]
    np.testing.assert_array_equal(m["euler_xyz"], r["euler_xyz"])
    a, _, c = r["euler_xyz"]
    Tx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    Tz = np.array([[math.cos(c), math.sin(c), 0], [-math.sin(c), math.cos(c), 0], [0, 0, 1]])
    np.testing.assert_allclose(m["T_rot"][:3, :3], Tx @ Tz, atol=2e-6)
