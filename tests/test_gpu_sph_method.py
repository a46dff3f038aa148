"""SphRegistration (PHZ_METHOD_SPH, CloudController("sph")) on the GPU
against the FP64 oracle of the same method, at the bars of the two-channel
parity tests: grids within 2e-3 of their maximum, argmax bins bit-exact
unless the oracle's own top values are within 1e-4 (near-tie), pose equal to
the oracle's closed form of the same bins.  The refine_rot_x/y/z masks leave
their Euler angles out of T_rot (and so out of the registered cloud and the
translation) as the oracle does; reflectivity and ambient (channels = 4) do
not change the rotation grid of this method; the C++ mirror's
CloudController("sph") equals the C ABI; both methods at the reference's
default sizes (bandwidth 75, 100 voxels)."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phaser_amd", "lib", "phaser_host_driver")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _argmax_ok(idx, grid):
    best = int(np.argmax(grid))
    return idx == best or grid[idx] >= grid[best] * (1 - 1e-4)


def _check(r, o):
    er, et = _rel(r["rot_corr"], o["rot_corr"]), _rel(r["trans_corr"], o["trans_corr"])
    print(f"  rot grid {er:.2e}, trans grid {et:.2e}, rot {r['rot_index']} vs {o['rot_index']}, "
          f"trans {r['trans_index']} vs {o['trans_index']}", flush=True)
    assert er < 2e-3 and et < 2e-3
    assert _argmax_ok(r["rot_index"], o["rot_corr"])
    if r["rot_index"] == o["rot_index"]:
        np.testing.assert_allclose(r["quat"], o["quat"], atol=1e-9)
        np.testing.assert_allclose(r["euler_xyz"], o["euler_xyz"], atol=1e-9)
        np.testing.assert_allclose(r["T_rot"], o["T_rot"], atol=2e-6)
        assert _argmax_ok(r["trans_index"], o["trans_corr"])
        if r["trans_index"] == o["trans_index"]:
            np.testing.assert_allclose(r["translation"], o["translation"], atol=1e-9)


@pytest.mark.parametrize("seed,skip", [(0, 0), (1, 0), (2, 5)])
def test_sph_method_matches_oracle(po, seed, skip):
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=seed, rings=64, azimuths=1024)
    print(f"sph method seed {seed} skip {skip}: {len(tgt)} / {len(src)} points", flush=True)
    with Context(bandwidth=32, n_voxels=64, max_points=200000, method=1,
                 skip_rot_axes=skip) as ctx:
        r = ctx.register(tgt, src, grids=True)
    o = po.register(tgt, src, bw=32, n_voxels=64, want_grids=True, method=1, skip_rot_axes=skip)
    _check(r, o)
    if skip:
        a, _, c = r["euler_xyz"]
        assert a != 0 or c != 0
        # x and z masked: T_rot = Ty of the same Euler angle
        np.testing.assert_allclose(np.asarray(r["T_rot"])[:3, :3], _ty(r), atol=2e-6)


def _ty(r):
    b = r["euler_xyz"][1]
    return np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])


def test_sph_rotation_ignores_reflectivity_and_ambient():
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=3, rings=64, azimuths=1024)
    e_t = np.stack([tgt[:, 3] * 2, tgt[:, 3] + 7], 1).astype(np.float32)
    e_s = np.stack([src[:, 3] * 2, src[:, 3] + 7], 1).astype(np.float32)
    print("sph method, four channels", flush=True)
    with Context(bandwidth=32, n_voxels=64, max_points=200000, method=1) as c2:
        a = c2.register(tgt, src, grids=True)
    with Context(bandwidth=32, n_voxels=64, max_points=200000, method=1, channels=4) as c4:
        b = c4.register(tgt, src, grids=True, target_extra=e_t, source_extra=e_s)
    assert a["rot_index"] == b["rot_index"]
    assert _rel(b["rot_corr"], a["rot_corr"]) < 1e-6
    np.testing.assert_array_equal(a["quat"], b["quat"])


def test_cloud_controller_sph_through_the_cpp_mirror(tmp_path):
    """host_driver with method "sph" and refine_rot_y false equals the C ABI
    with PHZ_METHOD_SPH and skip_rot_axes 2."""
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=4, rings=64, azimuths=1024)
    paths = []
    for name, cl in (("target", tgt), ("source", src)):
        p = tmp_path / f"{name}.bin"
        np.ascontiguousarray(cl, np.float32).tofile(p)
        paths.append(str(p))
    print("C++ driver, CloudController(\"sph\")", flush=True)
    r = subprocess.run([DRIVER, paths[0], paths[1], "32", "64", "0", "0", "sph", "101"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    with Context(bandwidth=32, n_voxels=64, max_points=262144, method=1, skip_rot_axes=2) as ctx:
        want = ctx.register(tgt, src)
    assert out["rot_index"] == want["rot_index"] and out["trans_index"] == want["trans_index"]
    np.testing.assert_allclose(out["state"][:4], want["quat"], atol=1e-12)
    np.testing.assert_allclose(out["state"][5:], want["translation"], atol=1e-9)


@pytest.mark.parametrize("method", [0, 1])
def test_reference_default_sizes_match_oracle(po, method):
    """The reference's default flags: bandwidth 75 (2B = 150, radix-5 SO(3)
    FFTs) and 100 voxels (radix 5 in every spatial pass), for both
    registration methods."""
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=7, rings=64, azimuths=1024)
    print(f"B=75 N=100 method {method}", flush=True)
    with Context(bandwidth=75, n_voxels=100, max_points=200000, method=method) as ctx:
        r = ctx.register(tgt, src, grids=True)
    o = po.register(tgt, src, bw=75, n_voxels=100, want_grids=True, method=method)
    _check(r, o)


def test_refine_flags_and_rotation_only_are_read_per_registration(tmp_path):
    """One CloudController("sph") registers twice: first with refine_rot_y
    false, then with every axis refined and FLAGS_estimate_translation false
    (sph-registration.cc:91-93, 108-113).  The second result is the
    rotation-only one: the C ABI's rotation with skip_rot_axes 0, no
    positional estimate, and the source rotated by T_rot but not moved."""
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=4, rings=64, azimuths=1024)
    paths = []
    for name, cl in (("target", tgt), ("source", src)):
        p = tmp_path / f"{name}.bin"
        np.ascontiguousarray(cl, np.float32).tofile(p)
        paths.append(str(p))
    print("C++ driver, CloudController(\"sph\"), two registrations", flush=True)
    r = subprocess.run([DRIVER, paths[0], paths[1], "32", "64", "0", "0", "sph", "101", "111",
                        "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    first, second = [json.loads(l) for l in r.stdout.strip().splitlines()[-2:]]
    with Context(bandwidth=32, n_voxels=64, max_points=262144, method=1, skip_rot_axes=2) as ctx:
        masked = ctx.register(tgt, src)
    with Context(bandwidth=32, n_voxels=64, max_points=262144, method=1) as ctx:
        full = ctx.register(tgt, src)
    assert first["rot_index"] == masked["rot_index"] and first["has_pos_estimate"] == 1
    np.testing.assert_allclose(first["state"][5:], masked["translation"], atol=1e-9)
    assert second["rot_index"] == full["rot_index"] and second["has_pos_estimate"] == 0
    np.testing.assert_allclose(second["state"][:4], full["quat"], atol=1e-12)
    assert second["state"][5:] == [0.0, 0.0, 0.0]
    T = np.asarray(full["T_rot"], np.float32).reshape(4, 4)
    p = np.ascontiguousarray(src[:, :3], np.float32)
    rot = np.stack([((T[i, 0] * p[:, 0] + T[i, 1] * p[:, 1]) + T[i, 2] * p[:, 2]) + T[i, 3]
                    for i in range(3)], 1).astype(np.float32)
    np.testing.assert_allclose(second["registered_sum"], float(rot.astype(np.float64).sum()),
                               rtol=1e-12)
