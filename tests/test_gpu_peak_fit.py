"""Peak-window fits on the device (bingham_peak_neighbors and
gaussian_peak_neighbors > 0) against the oracle's fits of the same grid
values: the GPU's own rotation and translation grids, returned by the same
call.  Both fits run in FP64 on identical inputs, so the mode quaternion,
omega, and the Gaussian mean and covariance agree to 1e-10 (the order of
the sums differs); euler_xyz is the oracle's closed form of the device's
mode to 1e-9 rad and T_rot its float transform to 2e-6.

Also: a call without output grids (the context's own buffers) gives the
same pose as one with them (to 1e-12 for the rotation; to 1e-5 m for the
translation, whose voxel sums are float atomics in arbitrary order); batched pairs each get the fit of their own
grids; 0 neighbours keep the argmax pose with omega (0, 0, 0, 1) and a zero
covariance; a Bingham window that ends at (2B)^3 fails with -4, as the
reference's CHECK_LT(end, n) (bingham-peak-based-eval.cc:69) aborts.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

CLOUDS = dict(rings=64, azimuths=1024)


def _pair(seed):
    from phaser_amd import synth
    return synth.pair(seed=seed, **CLOUDS)


def _check_fit(r, rot_corr, trans_corr, B, N, kb, kg):
    q, om = oracle.refine_rotation(B, np.asarray(rot_corr, np.float64), r["rot_index"], kb)
    print(f"  rot {r['rot_index']} quat {r['quat']} vs {q}; omega {r['rot_omega']}", flush=True)
    np.testing.assert_allclose(r["quat"], q, atol=1e-10)
    np.testing.assert_allclose(r["rot_omega"], om, atol=1e-10)
    A = r["rot_axes"]
    np.testing.assert_allclose(A.T @ A, np.eye(4), atol=1e-12)
    np.testing.assert_array_equal(A[:, 3], r["quat"])
    np.testing.assert_allclose(r["euler_xyz"], oracle.xyz_from_quat(r["quat"]), atol=1e-9)
    np.testing.assert_allclose(r["T_rot"], oracle.transform_xyz(*r["euler_xyz"]), atol=2e-6)
    mu, cov = oracle.refine_translation(N, np.asarray(trans_corr, np.float64), r["trans_index"], kg)
    print(f"  trans {r['trans_index']} mean {r['translation']} vs {mu}", flush=True)
    np.testing.assert_allclose(r["translation"], mu, atol=1e-9)
    np.testing.assert_allclose(r["trans_cov"], cov, atol=1e-9)


def _same_pose(a, b):
    """Two runs of one pair: the voxel sums are float atomics in arbitrary
    order, so the translation grid (and its fit) varies at FP32 rounding
    level between runs; the rotation side is deterministic."""
    assert a["rot_index"] == b["rot_index"] and a["trans_index"] == b["trans_index"]
    for key in ("quat", "rot_omega", "euler_xyz"):
        np.testing.assert_allclose(a[key], b[key], atol=1e-12, err_msg=key)
    for key in ("translation", "trans_cov"):
        np.testing.assert_allclose(a[key], b[key], atol=1e-5, err_msg=key)


@pytest.mark.parametrize("kb,kg", [(3, 3), (40, 25)])
def test_peak_fit_matches_oracle_on_device_grids(kb, kg):
    from phaser_amd.capi import Context
    B, N = 32, 64
    with Context(bandwidth=B, n_voxels=N, max_points=200000, bingham_peak_neighbors=kb,
                 gaussian_peak_neighbors=kg) as ctx:
        for seed in (0, 1):
            src, tgt, _, _ = _pair(seed)
            print(f"peak fit kb {kb} kg {kg} seed {seed}", flush=True)
            r = ctx.register(tgt, src, grids=True)
            _check_fit(r, r["rot_corr"], r["trans_corr"], B, N, kb, kg)
            q = ctx.register(tgt, src)  # the context's own grid buffers
            _same_pose(q, r)


def test_peak_fit_batched_pairs_use_their_own_grids():
    from phaser_amd.capi import Context
    B, N, kb, kg = 32, 64, 5, 6
    pairs = [_pair(s) for s in (2, 3, 4)]
    n3, T3 = (2 * B) ** 3, N ** 3
    with Context(bandwidth=B, n_voxels=N, max_points=200000, max_batch=4,
                 bingham_peak_neighbors=kb, gaussian_peak_neighbors=kg) as ctx:
        dt = [ctx.device_array(t) for _, t, _, _ in pairs]
        ds = [ctx.device_array(s) for s, _, _, _ in pairs]
        drot = ctx.device_empty(4 * n3 * len(pairs))
        dtr = ctx.device_empty(4 * T3 * len(pairs))
        print("peak fit batch of 3 with grids", flush=True)
        res = ctx.register_pairs_device([a.ptr for a in dt], [len(t) for _, t, _, _ in pairs],
                                        [a.ptr for a in ds], [len(s) for s, _, _, _ in pairs],
                                        d_rot=drot.ptr, d_trans=dtr.ptr)
        rot = drot.download((len(pairs), n3), np.float32)
        tr = dtr.download((len(pairs), T3), np.float32)
        for p, r in enumerate(res):
            _check_fit(r, rot[p], tr[p], B, N, kb, kg)
        print("peak fit batch of 3 without grids", flush=True)
        res2 = ctx.register_pairs_device([a.ptr for a in dt], [len(t) for _, t, _, _ in pairs],
                                         [a.ptr for a in ds], [len(s) for s, _, _, _ in pairs])
        for a, b in zip(res, res2):
            _same_pose(a, b)
        for a in dt + ds + [drot, dtr]:
            a.free()


def test_zero_neighbours_keep_the_argmax_pose():
    from phaser_amd.capi import Context
    B, N = 32, 64
    src, tgt, _, _ = _pair(5)
    with Context(bandwidth=B, n_voxels=N, max_points=200000) as ctx:
        r = ctx.register(tgt, src)
    np.testing.assert_allclose(r["quat"],  # device and host cos/sin may differ by an ulp
                               oracle.quat_from_zyz(oracle.zyz_from_index(r["rot_index"], B)),
                               atol=1e-15)
    np.testing.assert_array_equal(r["rot_omega"], [0, 0, 0, 1])
    A = r["rot_axes"]
    np.testing.assert_allclose(A.T @ A, np.eye(4), atol=1e-15)
    np.testing.assert_array_equal(A[:, 3], r["quat"])
    np.testing.assert_array_equal(r["trans_cov"], np.zeros((3, 3)))
    x, y, z = oracle.ind2sub(r["trans_index"], N, N)
    np.testing.assert_allclose(r["translation"],
                               [oracle.translation_from_index(v, N) for v in (x, y, z)], atol=1e-12)


def test_bingham_window_past_the_grid_fails():
    """k = (2B)^3 - argmax makes the window end at (2B)^3: the call returns
    -4 (the reference's CHECK_LT aborts); one neighbour fewer is a valid
    window, which ends at the grid's last index."""
    from phaser_amd.capi import Context, PhaserError
    B, N = 16, 32
    src, tgt, _, _ = _pair(6)
    with Context(bandwidth=B, n_voxels=N, max_points=200000) as ctx:
        idx = ctx.register(tgt, src)["rot_index"]
    k = (2 * B) ** 3 - idx
    print(f"argmax {idx}, bingham_peak_neighbors {k}", flush=True)
    with Context(bandwidth=B, n_voxels=N, max_points=200000, bingham_peak_neighbors=k) as ctx:
        with pytest.raises(PhaserError, match="error -4: pair 0: the Bingham peak window"):
            ctx.register(tgt, src)
        with pytest.raises(oracle.PeakWindowError):
            oracle.peak_window(idx, (2 * B) ** 3, k, "bingham")
        # -4 still fills every output: the result, both grids, the registered cloud
        failed = _raw_register(ctx, tgt, src, B, N)
    assert failed["rc"] == -4 and failed["status"] == 1
    with Context(bandwidth=B, n_voxels=N, max_points=200000,
                 bingham_peak_neighbors=k - 1) as ctx:
        r = ctx.register(tgt, src, grids=True, registered=True)
        assert r["rot_index"] == idx
        _check_fit(r, r["rot_corr"], r["trans_corr"], B, N, k - 1, 0)
    # the failed fit leaves the argmax pose: the outputs are those of a
    # registration without peak neighbours (its translation grid up to the
    # last bits of the FP32 atomic voxel sums, whose order varies)
    with Context(bandwidth=B, n_voxels=N, max_points=200000) as ctx:
        a = ctx.register(tgt, src, grids=True, registered=True)
    assert failed["rot_index"] == idx == a["rot_index"]
    np.testing.assert_array_equal(failed["rot_corr"], a["rot_corr"])
    np.testing.assert_array_equal(failed["quat"], a["quat"])
    np.testing.assert_allclose(failed["quat"], oracle.quat_from_zyz(oracle.zyz_from_index(idx, B)),
                               atol=1e-15)
    assert np.all(np.isfinite(failed["trans_corr"]))
    scale = np.max(np.abs(a["trans_corr"]))
    assert np.max(np.abs(failed["trans_corr"] - a["trans_corr"])) <= 1e-5 * scale
    assert failed["trans_index"] == a["trans_index"]
    np.testing.assert_array_equal(failed["translation"], a["translation"])
    assert np.all(np.isfinite(failed["registered"]))
    np.testing.assert_allclose(failed["registered"], a["registered"], atol=1e-5)
    np.testing.assert_array_equal(failed["registered"][:, 3], src[:, 3])


def _raw_register(ctx, tgt, src, B, N):
    """phz_register_channels through ctypes, keeping the outputs of a
    nonzero return."""
    import ctypes as C
    from phaser_amd.capi import Result, lib
    t = np.ascontiguousarray(tgt, np.float32)
    s = np.ascontiguousarray(src, np.float32)
    res = Result()
    rot = np.full((2 * B) ** 3, np.nan, np.float32)
    tr = np.full(N ** 3, np.nan, np.float32)
    reg = np.full_like(s, np.nan)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib().phz_register_channels(ctx.h, ptr(t), None, t.shape[0], ptr(s), None, s.shape[0],
                                     C.byref(res), ptr(rot), ptr(tr), ptr(reg))
    out = res.as_dict()
    out.update(rc=rc, rot_corr=rot, trans_corr=tr, registered=reg)
    return out
