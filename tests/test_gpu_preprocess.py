"""CloudPreProcessor's filters on the device (phz_voxel_grid_filter,
phz_pass_through_filter) against the oracle, bit-exact: same points, same
order, same float32 centroids (both sum each voxel's points in input order).
Clouds of 0 to 2,000,000 points, negative coordinates, non-finite points,
the overflow rule, the inclusive limits; then the C++ mirror
(preproc::CloudPreProcessor, and CloudController with the pre-processor
flags on) against the C ABI on the oracle-filtered clouds.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phaser_amd", "lib", "phaser_host_driver")


@pytest.fixture(scope="module")
def ctx():
    from phaser_amd.capi import Context
    with Context(bandwidth=8, n_voxels=8, max_points=16) as c:
        yield c


def _cloud(seed, n, lo, hi, holes=0):
    rng = np.random.default_rng(seed)
    p = rng.uniform(lo, hi, (n, 4)).astype(np.float32)
    if holes:
        i = rng.choice(n, holes, replace=False)
        p[i, rng.integers(0, 4, holes)] = rng.choice([np.nan, np.inf, -np.inf], holes)
    return p


@pytest.mark.parametrize("n,leaf,lo,hi,holes,seed", [
    (500, 0.25, 0.0, np.pi, 0, 0),            # the reference's test cloud
    (1, 0.25, -1.0, 1.0, 0, 1),
    (7, 0.5, -1.0, 1.0, 2, 2),
    (100000, 0.25, -50.0, 50.0, 100, 3),
    (2000000, 0.25, -60.0, 60.0, 1000, 4),    # a full driver-run scan
    (300000, 0.05, -3.0, 3.0, 0, 5),          # many points per voxel
    (50000, 2.0, -100.0, 20.0, 10, 6),
])
def test_voxel_grid_matches_oracle(po, ctx, n, leaf, lo, hi, holes, seed):
    p = _cloud(seed, n, lo, hi, holes)
    print(f"voxel grid n={n} leaf={leaf}", flush=True)
    got = ctx.voxel_grid_filter(p, leaf)
    want = po.voxel_grid_filter(p, leaf)
    print(f"  {len(want)} voxels", flush=True)
    assert got.shape == want.shape
    assert np.array_equal(got, want, equal_nan=True)


def test_voxel_grid_edge_cases(po, ctx):
    assert ctx.voxel_grid_filter(np.zeros((0, 4), np.float32)).shape == (0, 4)
    assert len(ctx.voxel_grid_filter(np.full((5, 4), np.nan, np.float32))) == 0
    p = _cloud(7, 100, -500, 500)
    p[3, 0] = np.nan
    assert np.array_equal(ctx.voxel_grid_filter(p, 1e-3), p, equal_nan=True)  # overflow
    for edge in (1289.5, 1290.5):  # either side of INT32_MAX voxels
        q = np.array([[0, 0, 0, 0], [edge, edge, edge, 1]], np.float32)
        assert np.array_equal(ctx.voxel_grid_filter(q, 1.0), po.voxel_grid_filter(q, 1.0))


@pytest.mark.parametrize("n,field,lo,hi,negative,seed", [
    (500, 2, -5.0, 3.0, True, 0),             # the command's defaults
    (2000000, 2, -5.0, 3.0, True, 1),
    (100000, 0, -1.0, 1.0, False, 2),
    (100000, 3, 0.5, 2.0, True, 3),
    (1, 2, -5.0, 3.0, True, 4),
])
def test_pass_through_matches_oracle(po, ctx, n, field, lo, hi, negative, seed):
    p = _cloud(seed, n, -8.0, 8.0, holes=min(n // 100, 500))
    p[: min(n, 20), field] = lo
    p[min(n, 20): min(n, 40), field] = hi
    got = ctx.pass_through_filter(p, field, lo, hi, negative)
    want = po.pass_through_filter(p, field, lo, hi, negative)
    assert np.array_equal(got, want, equal_nan=True)
    assert ctx.pass_through_filter(np.zeros((0, 4), np.float32)).shape == (0, 4)


@pytest.mark.parametrize("vg,pt", [(1, 0), (0, 1), (1, 1)])
def test_cpp_preprocessor_matches_oracle(tmp_path, po, vg, pt):
    p = _cloud(8, 200000, -20.0, 20.0, holes=50)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    p.tofile(fin)
    r = subprocess.run([DRIVER, "preprocess", str(fin), str(fout), str(vg), str(pt), "0.3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(fout, np.float32).reshape(-1, 4)
    want = p
    if vg:
        want = po.voxel_grid_filter(want, 0.3)
    if pt:
        want = po.pass_through_filter(want)
    assert json.loads(r.stdout)["points"] == len(want)
    assert np.array_equal(got, want, equal_nan=True)


def test_cpp_controller_with_voxel_grid_registers_the_filtered_clouds(tmp_path, po):
    """CloudController with phaser_pre_enable_voxel_grid_downsampling: the
    same registration as the C ABI on the oracle-filtered clouds."""
    from phaser_amd import synth
    from phaser_amd.capi import Context
    src, tgt, _, _ = synth.pair(seed=3, rings=64, azimuths=1024)
    tf, sf = tmp_path / "t.bin", tmp_path / "s.bin"
    np.asarray(tgt, np.float32).tofile(tf)
    np.asarray(src, np.float32).tofile(sf)
    env = dict(os.environ, PHZ_TEST_PRE="vg")
    r = subprocess.run([DRIVER, str(tf), str(sf), "32", "64"], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    tv, sv = po.voxel_grid_filter(tgt), po.voxel_grid_filter(src)
    assert len(tv) < len(tgt) and len(sv) < len(src)
    with Context(bandwidth=32, n_voxels=64, max_points=max(len(tv), len(sv))) as c:
        want = c.register(tv, sv)
    assert got["registered_points"] == len(sv)
    assert got["rot_index"] == want["rot_index"]
    assert got["trans_index"] == want["trans_index"]
