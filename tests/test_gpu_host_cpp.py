"""The C++ mirror of the reference API (phaser_core::CloudController("sph-opt")
over libphaser_host.so and the C ABI) registers a synthetic pair to the same
rotation bin and translation voxel as the FP64 oracle (near-ties excepted).
The driver runs as a child process, before this test process touches the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phaser_amd", "lib", "phaser_host_driver")


def _argmax_ok(idx, grid):
    best = int(np.argmax(grid))
    return idx == best or grid[idx] >= grid[best] * (1 - 1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 3])
def test_cpp_controller_matches_oracle(tmp_path, po, seed):
    from phaser_amd import synth
    src, tgt, _, _ = synth.pair(seed=seed, rings=64, azimuths=1024)
    tf, sf = tmp_path / "target.bin", tmp_path / "source.bin"
    tgt.astype(np.float32).tofile(tf)
    src.astype(np.float32).tofile(sf)
    print(f"host driver seed {seed}: {len(tgt)} / {len(src)} points", flush=True)
    r = subprocess.run([DRIVER, str(tf), str(sf), "32", "64"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    o = po.register(tgt, src, bw=32, n_voxels=64, want_grids=True)
    assert _argmax_ok(out["rot_index"], o["rot_corr"])
    if out["rot_index"] == o["rot_index"]:
        np.testing.assert_allclose(np.abs(out["state"][:4]), np.abs(o["quat"]), atol=1e-9)
        assert _argmax_ok(out["trans_index"], o["trans_corr"])
        if out["trans_index"] == o["trans_index"]:
            np.testing.assert_allclose(out["state"][5:], o["translation"], atol=1e-9)
    assert out["registered_points"] == len(src)
    assert out["rot_corr_size"] == 64 ** 3


@pytest.mark.gpu
def test_cpp_controller_peak_neighbours_match_c_abi(tmp_path):
    """FLAGS_bingham_peak_neighbors / FLAGS_gaussian_peak_neighbors reach the
    context: the C++ result (pose, Bingham omega, Gaussian covariance)
    equals the C ABI's for the same flags (translation to 1e-5: the voxel
    sums are float atomics in arbitrary order)."""
    from phaser_amd import synth
    src, tgt, _, _ = synth.pair(seed=1, rings=64, azimuths=1024)
    tf, sf = tmp_path / "target.bin", tmp_path / "source.bin"
    tgt.astype(np.float32).tofile(tf)
    src.astype(np.float32).tofile(sf)
    print("host driver with peak neighbours 4 / 5", flush=True)
    r = subprocess.run([DRIVER, str(tf), str(sf), "32", "64", "4", "5"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    from phaser_amd.capi import Context
    with Context(bandwidth=32, n_voxels=64, max_points=262144, bingham_peak_neighbors=4,
                 gaussian_peak_neighbors=5) as ctx:
        want = ctx.register(tgt, src)
    assert out["rot_index"] == want["rot_index"] and out["trans_index"] == want["trans_index"]
    np.testing.assert_allclose(out["state"][:4], want["quat"], atol=1e-12)
    np.testing.assert_allclose(out["rot_omega"], want["rot_omega"], atol=1e-12)
    assert want["rot_omega"][2] > 0  # a real multi-sample fit
    np.testing.assert_allclose(out["state"][5:], want["translation"], atol=1e-5)
    np.testing.assert_allclose(out["trans_cov_diag"], np.diag(want["trans_cov"]), atol=1e-5)


def test_cpp_driver_is_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    assert os.path.exists(os.path.join(ROOT, "phaser_amd", "lib", "libphaser_host.so"))
