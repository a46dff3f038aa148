"""Pins of the oracle's CloudPreProcessor filters (oracle.voxel_grid_filter,
oracle.pass_through_filter): the reference's own tests
(cloud-pre-processor-test.cc:55-81: 500 points uniform in [0, pi)^4, both
commands return fewer points), an independent brute force of PCL 1.8's
VoxelGrid (voxels as (floor(x/l), floor(y/l), floor(z/l)) tuples in float32,
ordered z, then y, then x; FP64 means), closed forms (one point per voxel:
the input permuted, copies of one point: that point, two points: their float
mean), the overflow rule, non-finite points, and PassThrough's branches
point by point (passthrough.hpp: inclusive limits, negative, non-finite).
"""
import numpy as np
import pytest


def _cloud(rng, n, lo=0.0, hi=np.pi):
    return rng.uniform(lo, hi, (n, 4)).astype(np.float32)


def _brute_voxel_grid(pts, leaf):
    inv = np.float32(1.0) / np.float32(leaf)
    groups = {}
    for p in pts:
        if not np.all(np.isfinite(p[:3])):
            continue
        key = tuple(int(np.floor(np.float32(p[k]) * inv)) for k in (2, 1, 0))
        groups.setdefault(key, []).append(p.astype(np.float64))
    return [np.mean(groups[k], axis=0) for k in sorted(groups)], [len(groups[k]) for k in sorted(groups)]


def test_reference_voxel_grid_test_property(po):
    """VoxelGridCmdTest: 500 points in [0, pi)^4 at leaf 0.25 -> fewer."""
    rng = np.random.default_rng(0)
    for _ in range(5):
        p = _cloud(rng, 500)
        assert 0 < len(po.voxel_grid_filter(p, 0.25)) < 500


def test_reference_pass_through_test_property(po):
    """PassThroughGndFilterTest: z in [-5, 3] removed -> fewer (only the
    points with z in (3, pi) stay)."""
    rng = np.random.default_rng(1)
    p = _cloud(rng, 500)
    out = po.pass_through_filter(p)
    assert len(out) < 500
    assert np.all(out[:, 2] > 3.0)
    assert len(out) == int(np.sum(p[:, 2] > np.float32(3.0)))


@pytest.mark.parametrize("leaf,lo,hi,n,seed", [(0.25, 0.0, np.pi, 500, 2), (0.1, -3.0, 2.0, 4000, 3),
                                               (1.0, -50.0, 50.0, 20000, 4),
                                               (0.3, -0.7, 0.7, 3000, 5)])
def test_voxel_grid_matches_brute_force(po, leaf, lo, hi, n, seed):
    """Grouping by the float32 floor of p / leaf per axis, the output
    ordered by (z, y, x) voxel -- PCL's linear index with x fastest -- and
    the centroid of every field within FP32 rounding of the FP64 mean."""
    rng = np.random.default_rng(seed)
    p = _cloud(rng, n, lo, hi)
    out = po.voxel_grid_filter(p, leaf)
    means, counts = _brute_voxel_grid(p, leaf)
    assert len(out) == len(means)
    assert max(counts) > 2
    np.testing.assert_allclose(out, np.array(means), rtol=2e-6, atol=2e-6 * max(abs(lo), hi))


def test_one_point_per_voxel_is_a_permutation(po):
    """Points at distinct voxel centres come back unchanged, sorted by
    voxel (x fastest)."""
    leaf = 0.5
    g = np.stack(np.meshgrid(np.arange(-3, 4), np.arange(-2, 3), np.arange(0, 3), indexing="ij"),
                 axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(6)
    g = g[rng.permutation(len(g))]
    p = np.concatenate([(g + 0.5) * leaf, rng.uniform(0, 1, (len(g), 1))], axis=1).astype(np.float32)
    out = po.voxel_grid_filter(p, leaf)
    order = np.lexsort((g[:, 0], g[:, 1], g[:, 2]))
    assert np.array_equal(out, p[order])


def test_copies_and_pairs_closed_forms(po):
    # short mantissas: the sum of seven copies is exact, so is the mean
    a = np.array([[1.25, -2.5, 0.625, 5.0]], np.float32)
    assert np.array_equal(po.voxel_grid_filter(np.repeat(a, 7, axis=0), 0.25), a)
    a = np.array([[1.3, -2.2, 0.7, 5.0]], np.float32)
    b = a + np.float32(0.01)
    out = po.voxel_grid_filter(np.concatenate([a, b]), 0.25)
    assert np.array_equal(out, ((a + b) / np.float32(2)).astype(np.float32))


def test_floor_not_truncation_near_zero(po):
    """-0.1 and +0.1 fall into voxels -1 and 0 at leaf 0.25 (floor), so two
    centroids, the negative one first."""
    p = np.array([[-0.1, 0, 0, 1], [0.1, 0, 0, 2]], np.float32)
    out = po.voxel_grid_filter(p, 0.25)
    assert np.array_equal(out, p)


def test_overflow_leaves_the_cloud_unchanged(po):
    """More than INT32_MAX voxels (voxel_grid.hpp's overflow check): the
    input, non-finite points included."""
    rng = np.random.default_rng(7)
    p = _cloud(rng, 100, -500, 500)
    p[3, 0] = np.nan
    out = po.voxel_grid_filter(p, 1e-3)
    assert np.array_equal(out, p, equal_nan=True)
    # 1290^3 < 2^31 - 1 < 1291^3 voxels at the edge of the check
    q = np.array([[0, 0, 0, 0], [1289.5, 1289.5, 1289.5, 0]], np.float32)
    assert len(po.voxel_grid_filter(q, 1.0)) == 2
    q[1, :3] = 1290.5
    assert np.array_equal(po.voxel_grid_filter(q, 1.0), q)


def test_non_finite_points_are_dropped(po):
    rng = np.random.default_rng(8)
    p = _cloud(rng, 200)
    p[5, 1] = np.inf
    p[9, 2] = np.nan
    p[11, 3] = np.nan  # a non-finite intensity does not drop the point
    out = po.voxel_grid_filter(p, 0.25)
    fin = np.delete(p, [5, 9], axis=0)
    ref = po.voxel_grid_filter(fin, 0.25)
    assert np.array_equal(out, ref, equal_nan=True)
    assert np.isnan(out[:, 3]).sum() == 1
    assert len(po.voxel_grid_filter(np.full((3, 4), np.nan, np.float32), 0.25)) == 0


def _brute_pass_through(p, field, lo, hi, negative):
    keep = []
    for q in p:
        if not np.all(np.isfinite(q[:3])):
            continue
        v = q[field]
        if not np.isfinite(v):
            continue
        if not negative and (v < lo or v > hi):
            continue
        if negative and lo <= v <= hi:
            continue
        keep.append(q)
    return np.array(keep, np.float32).reshape(-1, 4)


@pytest.mark.parametrize("field,lo,hi,negative", [(2, -5.0, 3.0, True), (2, -5.0, 3.0, False),
                                                  (0, -1.0, 1.0, False), (3, 0.5, 2.0, True)])
def test_pass_through_branches(po, field, lo, hi, negative):
    rng = np.random.default_rng(9)
    p = _cloud(rng, 2000, -8.0, 8.0)
    p[:20, field] = lo  # the inclusive limits
    p[20:40, field] = hi
    p[40, 0] = np.nan
    p[41, 3] = np.inf
    out = po.pass_through_filter(p, field, lo, hi, negative)
    assert np.array_equal(out, _brute_pass_through(p, field, lo, hi, negative))
    inside = (out[:, field] >= lo) & (out[:, field] <= hi)
    assert np.all(~inside if negative else inside)
