"""The peak-window fits (Bingham fit of the rotation and Gaussian fit of the
translation to the correlation values around each argmax) on the CPU.

* The oracle's fits against the reference's own known-answer vectors
  (tests/golden/peak_fit_kat.json, from bingham-test.cc and
  gaussian-test.cc), at 1e-9 (the reference tests at 1e-4).
* The window rules of bingham-peak-based-eval.cc:47-58 and
  gaussian-peak-based-eval.cc:26-37, case by case.
* Closed forms: two equally weighted quaternions; a window of symmetric
  weights along the gamma axis (the mode is the peak's quaternion and the
  two nonzero eigenvalues are sum w cos^2 and sum w sin^2 of the half angle
  offsets); a symmetric window along y (the mean is the peak's translation,
  the variance sum w dy^2); zero neighbours reduce to the argmax.
* The device's pose math (phaser_amd/csrc/pose.hpp, built for the host with
  hipcc) against numpy and the oracle.
"""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = json.load(open(os.path.join(ROOT, "tests", "golden", "peak_fit_kat.json")))
D = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(D)


def test_bingham_mode_matches_reference_kats():
    for case in KAT["bingham"]:
        m = oracle.bingham_mode(case["samples"], case["weights"])
        np.testing.assert_allclose(m, case["mode"], atol=1e-9, err_msg=case["name"])


def test_gaussian_fit_matches_reference_kat():
    case = KAT["gaussian"][0]
    mu, cov = oracle.gaussian_fit(case["samples"], case["weights"])
    np.testing.assert_allclose(mu, case["mean"], atol=1e-12)
    np.testing.assert_allclose(cov, case["cov"], atol=1e-12)


def test_gaussian_fit_negative_weight_is_nan():
    """gaussian.cc:52-54 scales by sqrt(w): a negative weight poisons the
    covariance (the mean stays finite)."""
    mu, cov = oracle.gaussian_fit(np.arange(1, 10.0).reshape(3, 3), [-0.1, 0.6, 0.5])
    assert np.all(np.isfinite(mu)) and np.all(np.isnan(cov))


def test_bingham_two_sample_closed_form():
    """w = (1/2, 1/2): C = (q1 q1^T + q2 q2^T) / 2 has eigenvalues
    (1 +- |c|) / 2, c = q1 . q2, with the mode (q1 + sign(c) q2) / norm."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        q1, q2 = rng.normal(size=(2, 4))
        q1 /= np.linalg.norm(q1)
        q2 /= np.linalg.norm(q2)
        c = q1 @ q2
        want = q1 + np.sign(c) * q2
        want /= np.linalg.norm(want)
        got = oracle.bingham_mode(np.stack([q1, q2], 1), [0.5, 0.5], ref=q1)
        np.testing.assert_allclose(got, want, atol=1e-12)
        ev = np.linalg.eigvalsh(oracle.bingham_moment(np.stack([q1, q2], 1), [0.5, 0.5]))
        np.testing.assert_allclose(ev[2:], [(1 - abs(c)) / 2, (1 + abs(c)) / 2], atol=1e-12)


def test_peak_window_rules():
    n = 100
    w = oracle.peak_window
    for kind in ("bingham", "gaussian"):
        assert w(50, n, 3, kind) == (47, 53)
        assert w(2, n, 3, kind) == (2, 5)  # k > idx: no lower neighbours at all
        assert w(3, n, 3, kind) == (0, 6)
        assert w(50, n, 0, kind) == (50, 50)
        assert w(96, n, 3, kind) == (93, 99)
    assert w(98, n, 3, "bingham") == (95, 98)  # idx + k > n: end = idx
    with pytest.raises(oracle.PeakWindowError):  # idx + k == n: end = n fails CHECK_LT
        w(97, n, 3, "bingham")
    assert w(97, n, 3, "gaussian") == (94, 97)  # idx + k >= n: end = idx
    assert w(99, n, 200, "gaussian") == (99, 99)
    assert w(99, n, 200, "bingham") == (99, 99)


def test_zero_neighbours_reduce_to_the_argmax():
    bw, N = 4, 8
    rng = np.random.default_rng(0)
    corr = rng.random((2 * bw) ** 3)
    idx = int(np.argmax(corr))
    q, om = oracle.refine_rotation(bw, corr, idx, 0)
    np.testing.assert_array_equal(q, oracle.quat_from_zyz(oracle.zyz_from_index(idx, bw)))
    np.testing.assert_array_equal(om, [0, 0, 0, 1])
    corr3 = rng.random(N ** 3)
    j = int(np.argmax(corr3))
    mu, cov = oracle.refine_translation(N, corr3, j, 0)
    x, y, z = oracle.ind2sub(j, N, N)
    np.testing.assert_array_equal(mu, [oracle.translation_from_index(v, N) for v in (x, y, z)])
    np.testing.assert_array_equal(cov, np.zeros((3, 3)))
    with pytest.raises(oracle.PeakWindowError):  # CHECK_GT(sum, 0) holds at 0 neighbours too
        oracle.refine_translation(N, -corr3, j, 0)


def _gamma_window(bw, k, seed):
    """A grid whose window around the peak lies inside one gamma row, with
    weights symmetric about it."""
    n = 2 * bw
    rng = np.random.default_rng(seed)
    corr = rng.random(n ** 3) * 0.1
    ii, jj, kk = 3, 5, n // 2
    idx = (ii * n + jj) * n + kk
    half = rng.random(k) * 0.5 + 0.2
    corr[idx] = 2.0
    corr[idx + 1:idx + k + 1] = half
    corr[idx - k:idx][::-1] = half
    return corr, idx


@pytest.mark.parametrize("k", [1, 3, 6])
def test_bingham_symmetric_gamma_window_closed_form(k):
    """Along gamma q(g) = q_ab (cos g/2, 0, 0, sin g/2): with weights
    symmetric about the peak g0 the mode is q(g0), and omega's two nonzero
    entries are sum w sin^2(d/2) and sum w cos^2(d/2), d = g - g0."""
    bw = 8
    corr, idx = _gamma_window(bw, k, k)
    q, om = oracle.refine_rotation(bw, corr, idx, k)
    np.testing.assert_allclose(q, oracle.quat_from_zyz(oracle.zyz_from_index(idx, bw)), atol=1e-12)
    w = corr[idx - k:idx + k + 1] / corr[idx - k:idx + k + 1].sum()
    d = np.pi * np.arange(-k, k + 1) / bw  # gamma step pi / bw
    np.testing.assert_allclose(om, [0, 0, w @ np.sin(d / 2) ** 2, w @ np.cos(d / 2) ** 2],
                               atol=1e-12)


@pytest.mark.parametrize("k", [1, 4])
def test_gaussian_symmetric_y_window_closed_form(k):
    """Consecutive linear indices step y by one voxel: with weights symmetric
    about the peak (away from the wrap at N/2) the mean is the peak's
    translation and the covariance is diag(0, sum w dy^2, 0)."""
    N = 32
    rng = np.random.default_rng(k)
    corr = rng.random(N ** 3) * 0.1
    x, y, z = 3, 6, 2
    idx = oracle.sub2ind(x, y, z, N, N)
    assert list(oracle.ind2sub(idx, N, N)) == [x, y, z]
    half = rng.random(k) + 0.2
    corr[idx] = 3.0
    corr[idx + 1:idx + k + 1] = half
    corr[idx - k:idx][::-1] = half
    mu, cov = oracle.refine_translation(N, corr, idx, k)
    np.testing.assert_allclose(mu, [oracle.translation_from_index(v, N) for v in (x, y, z)],
                               atol=1e-12)
    w = corr[idx - k:idx + k + 1] / corr[idx - k:idx + k + 1].sum()
    step = 101.0 / N  # (|-50| + |50| + 1) / N
    want = np.zeros((3, 3))
    want[1, 1] = w @ (step * np.arange(-k, k + 1)) ** 2
    np.testing.assert_allclose(cov, want, atol=1e-12)


# ---- the device's pose math, built for the host ----------------------------

@pytest.fixture(scope="module")
def pose_lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    out = tmp_path_factory.mktemp("pose") / "libpose_host.so"
    subprocess.run([hipcc, "-O2", "-x", "hip", "--offload-host-only", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "phaser_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "pose_host.cc"), "-o", str(out)],
                   check=True, capture_output=True)
    return C.CDLL(str(out))


def test_eigen_sym4_matches_numpy(pose_lib):
    rng = np.random.default_rng(1)
    for trial in range(200):
        X = rng.normal(size=(4, 4))
        if trial % 4 == 1:  # near rank one, as moment matrices of tight windows are
            v = rng.normal(size=4)
            X = np.outer(v, v) + 1e-9 * X
        A = 0.5 * (X + X.T)
        a, d, V = A.copy(), np.zeros(4), np.zeros((4, 4))
        pose_lib.pose_eigen_sym4(_p(a), _p(d), _p(V))
        ev = np.linalg.eigvalsh(A)
        scale = np.abs(ev).max()
        np.testing.assert_allclose(d, ev, atol=1e-13 * scale)
        np.testing.assert_allclose(V.T @ V, np.eye(4), atol=1e-13)
        np.testing.assert_allclose(A @ V, V * d[None, :], atol=1e-12 * scale)


def test_bingham_from_moment_matches_kats_and_oracle(pose_lib):
    for case in KAT["bingham"][:2]:
        S = np.asarray(case["samples"])
        Cm = oracle.bingham_moment(S, case["weights"])
        tri = np.array([Cm[r, c] for r in range(4) for c in range(r, 4)])
        qp = S[:, 0].copy()
        mode, om, ax = np.zeros(4), np.zeros(4), np.zeros(16)
        pose_lib.pose_bingham_from_moment(_p(tri), _p(qp), 1, _p(mode), _p(om), _p(ax))
        np.testing.assert_allclose(mode, case["mode"], atol=1e-9)
        np.testing.assert_allclose(mode, oracle.bingham_mode(S, case["weights"], ref=qp),
                                   atol=1e-13)
        ev = np.linalg.eigvalsh(Cm)
        np.testing.assert_allclose(om, ev / ev.sum(), atol=1e-13)
        A = ax.reshape(4, 4)
        np.testing.assert_allclose(A[:, 3], mode, atol=0)
        np.testing.assert_allclose(A.T @ A, np.eye(4), atol=1e-13)
    # one sample: mode = the peak quaternion, omega (0, 0, 0, 1), orthonormal axes
    qp = np.asarray(KAT["bingham"][2]["mode"])
    qp = qp / np.linalg.norm(qp)
    mode, om, ax = np.zeros(4), np.zeros(4), np.zeros(16)
    pose_lib.pose_bingham_from_moment(_p(np.zeros(10)), _p(qp), 0, _p(mode), _p(om), _p(ax))
    np.testing.assert_array_equal(mode, qp)
    np.testing.assert_array_equal(om, [0, 0, 0, 1])
    A = ax.reshape(4, 4)
    np.testing.assert_allclose(A.T @ A, np.eye(4), atol=1e-15)
    np.testing.assert_array_equal(A[:, 3], qp)


def test_index_maps_and_euler_match_oracle(pose_lib):
    rng = np.random.default_rng(2)
    q, t, e, T = np.zeros(4), np.zeros(3), np.zeros(3), np.zeros(16, np.float32)
    for bw in (4, 16, 64):
        for idx in rng.integers(0, (2 * bw) ** 3, 20):
            pose_lib.pose_quat_of_index(C.c_uint32(int(idx)), bw, _p(q))
            want = oracle.quat_from_zyz(oracle.zyz_from_index(int(idx), bw))
            np.testing.assert_allclose(q, want, atol=1e-15)
            pose_lib.pose_euler_and_transform(_p(q), _p(e), T.ctypes.data_as(C.c_void_p))
            np.testing.assert_allclose(e, oracle.xyz_from_quat(q), atol=1e-14)
    for N in (8, 64, 128):
        for idx in rng.integers(0, N ** 3, 20):
            pose_lib.pose_translation_of_index(C.c_uint32(int(idx)), N, C.c_double(101.0), _p(t))
            x, y, z = oracle.ind2sub(int(idx), N, N)
            np.testing.assert_array_equal(t, [oracle.translation_from_index(v, N)
                                              for v in (x, y, z)])
