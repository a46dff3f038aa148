"""The mixture evaluations and the z-score peak extraction on the device,
against the oracle on the device's own grids (FP64 on identical inputs):

* rot_eval = PHZ_EVAL_MIXTURE (BmmPeakBasedEval, bmm_peak_neighbors) and
  pos_eval = PHZ_EVAL_MIXTURE (GmmPeakBasedEval, gmm_peak_neighbors) with the
  single argmax signal of ZScoreEval: quaternion and omega of the BMM
  window's Bingham, mean and covariance of the GMM window's Gaussian
  (1e-10), and the GMM window's CHECK_LT failure as status
  PHZ_PAIR_TRANS_WINDOW (-4, every result filled);
* phz_zscore_peaks (ZScorePeakExtraction) on hand-worked inputs, random
  spiky vectors and whole correlation grids: the same signal list as the
  oracle's detector run on the same float values.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    from phaser_amd import synth
    return synth.pair(seed=4, rings=64, azimuths=1024)


@pytest.mark.parametrize("kb,kg", [(2, 2), (5, 3), (0, 0)])
def test_mixture_evaluations_match_oracle(po, pair, kb, kg):
    from phaser_amd.capi import Context
    src, tgt, _, _ = pair
    B, N = 32, 64
    with Context(bandwidth=B, n_voxels=N, rot_eval=1, pos_eval=1, bmm_peak_neighbors=kb,
                 gmm_peak_neighbors=kg, max_points=max(len(src), len(tgt))) as c:
        r = c.register(tgt, src, grids=True)
    assert r["status"] == 0
    rc = r["rot_corr"].astype(np.float64)
    tc = r["trans_corr"].astype(np.float64)
    q, omega = po.refine_rotation(B, rc, r["rot_index"], kb, kind="bmm")
    np.testing.assert_allclose(r["quat"], q, atol=1e-10)
    np.testing.assert_allclose(r["rot_omega"], omega, atol=1e-10)
    mu, cov = po.refine_translation_gmm(N, tc, [r["trans_index"]], kg)
    np.testing.assert_allclose(r["translation"], mu, atol=1e-10)
    np.testing.assert_allclose(r["trans_cov"], cov, atol=1e-10)


def test_gmm_window_past_the_grid_fails_its_check(pair):
    """gmm_peak_neighbors = N^3 - argmax: the window's end would be N^3
    (gmm-peak-based-eval.cc:63-67), the reference's CHECK_LT(end, n) aborts
    (:81); here -4 with PHZ_PAIR_TRANS_WINDOW, every output filled."""
    from phaser_amd.capi import Context, PhaserError
    src, tgt, _, _ = pair
    with Context(bandwidth=32, n_voxels=64, max_points=max(len(src), len(tgt))) as c:
        idx = c.register(tgt, src)["trans_index"]
    k = 64 ** 3 - idx
    with Context(bandwidth=32, n_voxels=64, pos_eval=1, gmm_peak_neighbors=k,
                 max_points=max(len(src), len(tgt))) as c:
        with pytest.raises(PhaserError) as e:
            c.register(tgt, src)
    assert "gmm-peak-based-eval.cc:81" in str(e.value)


def _hand_cases():
    Z = dict(lag_percentile=0.1, lag_max=3, threshold=4.33, influence=0.05)
    ones = np.ones(30)

    def spikes(**kv):
        x = ones.copy()
        for k, v in kv.items():
            x[int(k[1:])] = v
        return x
    return [(spikes(i10=10.0), dict(Z, filter_threshold=0.0)),
            (spikes(i10=10.0, i11=1.2), dict(Z, filter_threshold=0.0)),
            (spikes(i10=10.0, i11=10.0, i14=3.5), dict(Z, filter_threshold=0.0)),
            (spikes(i10=-5.0, i12=1.7), dict(Z, filter_threshold=-100.0)),
            (spikes(i6=1.6, i10=0.5, i11=1.05), dict(Z, filter_threshold=0.9))]


def test_zscore_hand_cases_match_oracle(po):
    from phaser_amd.capi import Context
    with Context(bandwidth=8, n_voxels=8, max_points=16) as c:
        for x, kw in _hand_cases():
            got = c.zscore_peaks(x, **kw)
            want = po.zscore_peaks(x.astype(np.float32).astype(np.float64), **kw)
            assert got.tolist() == want.tolist(), (got, want)


@pytest.mark.parametrize("n,lag_max,filt,seed", [(1000, 20, 0.405, 0), (5000, 7, 0.0, 1),
                                                 (70000, 64, 0.2, 2), (257, 3, -1.0, 3)])
def test_zscore_random_vectors_match_oracle(po, n, lag_max, filt, seed):
    """Random vectors with clusters of spikes (replacements that feed one
    another within a 64-index step) and runs below the filter."""
    from phaser_amd.capi import Context
    rng = np.random.default_rng(seed)
    x = rng.gamma(2.0, 0.5, n)
    for c0 in rng.choice(n - 8, size=max(1, n // 50), replace=False):
        x[c0:c0 + rng.integers(1, 8)] += rng.uniform(2, 20)
    x[rng.choice(n, size=n // 10, replace=False)] = rng.uniform(-1, 0.3, n // 10)
    x = x.astype(np.float32)
    with Context(bandwidth=8, n_voxels=8, max_points=16) as c:
        got = c.zscore_peaks(x, lag_max=lag_max, filter_threshold=filt)
    want = po.zscore_peaks(x.astype(np.float64), lag_max=lag_max, filter_threshold=filt)
    print(f"z-score n={n} lag_max={lag_max}: {len(want)} signals", flush=True)
    assert len(want) > 0
    assert got.tolist() == want.tolist()


def test_zscore_on_correlation_grids_matches_oracle(po, pair):
    """Both grids of a registration (262,144 entries each at B = 32 / 64^3)
    with the flags' defaults."""
    from phaser_amd.capi import Context
    src, tgt, _, _ = pair
    with Context(bandwidth=32, n_voxels=64, max_points=max(len(src), len(tgt))) as c:
        r = c.register(tgt, src, grids=True)
        for g in (r["rot_corr"], r["trans_corr"]):
            got = c.zscore_peaks(g)
            want = po.zscore_peaks(g.astype(np.float64))
            print(f"z-score grid: {len(want)} signals", flush=True)
            assert got.tolist() == want.tolist()
