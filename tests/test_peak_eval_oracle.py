"""Pins of the oracle's mixture evaluations and z-score peak extraction
(CPU): the reference's known answers, the window rules case by case, and
hand-worked runs of the z-score recurrence.

* GaussianMixture::calcMixutreParameters: gaussian-mixture-test.cc:15-58
  (CalcGMComponentsTest, CalcGMComponentsFromGaussiansTest).
* GmmPeakBasedEval's window (gmm-peak-based-eval.cc:58-68, CHECK_LT :81)
  against GaussianPeakBasedEval's (gaussian-peak-based-eval.cc:26-37).
* BmmPeakBasedEval normalises one sample too (bmm-peak-based-eval.cc:83-85).
* ZScorePeakExtraction (z-score-peak-extraction.cc:42-112): small inputs
  whose signals follow by hand from the reference's loop, chosen so that a
  dropped replacement, a window that includes the current index, a filter
  that still updates the window, or signals below the mean each change the
  answer.
"""
import numpy as np
import pytest


def test_gaussian_mixture_reference_kats(po):
    # gaussian-mixture-test.cc:15-31 (2-D means as columns, identity covs)
    means = np.array([[1.0, 2.0], [3.0, 4.0]])
    mu, cov = po.gaussian_mixture(means, [np.eye(2), np.eye(2)], np.array([0.4, 0.6]))
    np.testing.assert_allclose(mu, [1.6, 3.6], atol=1e-4)
    np.testing.assert_allclose(cov, [[1.24, 0.24], [0.24, 1.24]], atol=1e-4)
    # gaussian-mixture-test.cc:33-56
    means = np.array([[1.0, 2.0], [2.0, 2.0], [3.0, 1.0]])
    covs = [np.diag([1.0, 2.0, 3.0]), np.diag([4.0, 5.0, 6.0])]
    mu, cov = po.gaussian_mixture(means, covs, np.array([0.3, 0.7]))
    np.testing.assert_allclose(mu, [1.7, 2.0, 1.6], atol=1e-4)
    np.testing.assert_allclose(cov, [[3.31, 0, -0.42], [0, 4.1, 0], [-0.42, 0, 5.94]], atol=1e-4)


def test_mixture_window_rules(po):
    n = 100
    # the Gaussian window collapses to the peak at idx + k >= n, the GMM's
    # only past n, and an end of n fails the GMM's CHECK_LT
    assert po.peak_window(97, n, 3, "gaussian") == (94, 97)
    with pytest.raises(po.PeakWindowError):
        po.peak_window(97, n, 3, "gmm")
    assert po.peak_window(98, n, 3, "gmm") == (95, 98)
    assert po.peak_window(96, n, 3, "gmm") == (93, 99)
    assert po.peak_window(2, n, 3, "gmm") == (2, 5)
    # the BMM window is the Bingham one
    for idx in (0, 2, 50, 96, 98):
        try:
            want = po.peak_window(idx, n, 3, "bingham")
        except po.PeakWindowError:
            with pytest.raises(po.PeakWindowError):
                po.peak_window(idx, n, 3, "bmm")
            continue
        assert po.peak_window(idx, n, 3, "bmm") == want


def test_bmm_normalises_one_sample(po):
    corr = np.zeros(8 ** 3)
    corr[10] = 2.5
    s, w, _ = po.rotation_window(4, corr, 10, 0, "bmm")
    assert s.shape == (4, 1) and w.tolist() == [1.0]
    s, w, _ = po.rotation_window(4, corr, 10, 0, "bingham")
    assert w.tolist() == [2.5]  # the Bingham evaluation leaves one sample as it is
    corr[10] = -1.0
    with pytest.raises(po.PeakWindowError):
        po.rotation_window(4, corr, 10, 0, "bmm")


def test_gmm_of_one_signal_is_its_window_gaussian(po):
    rng = np.random.default_rng(3)
    N = 8
    corr = rng.uniform(0.1, 1.0, N ** 3)
    idx = 200
    mu, cov = po.refine_translation_gmm(N, corr, [idx], 2)
    m1, c1 = po.refine_translation(N, corr, idx, 2, kind="gmm")
    np.testing.assert_array_equal(mu, m1)
    np.testing.assert_allclose(cov, c1, atol=1e-15)
    # two signals: the mixture of the two window Gaussians
    mu2, cov2 = po.refine_translation_gmm(N, corr, [idx, 300], 2)
    m2, c2 = po.refine_translation(N, corr, 300, 2, kind="gmm")
    w = np.array([corr[idx], corr[300]]) / (corr[idx] + corr[300])
    np.testing.assert_allclose(mu2, w[0] * m1 + w[1] * m2, atol=1e-12)
    d1, d2 = m1 - mu2, m2 - mu2
    want = w[0] * c1 + w[1] * c2 + w[0] * np.outer(d1, d1) + w[1] * np.outer(d2, d2)
    np.testing.assert_allclose(cov2, want, atol=1e-12)


Z = dict(lag_percentile=0.1, lag_max=3, threshold=4.33, influence=0.05)


def _ones(n=30, **spikes):
    x = np.ones(n)
    for k, v in spikes.items():
        x[int(k[1:])] = v
    return x


def test_zscore_single_spike(po):
    # lag 3: windows of three ones have deviation 0, so the spike at 10 is a
    # signal; it becomes 0.05 * 10 + 0.95 * 1 = 1.45, and the later windows
    # [8, 11), [9, 12), [10, 13) hold it (mean 1.15, deviation 0.26): the
    # ones after it stay within 4.33 deviations
    assert po.zscore_peaks(_ones(i10=10.0), filter_threshold=0.0, **Z).tolist() == [10]


def test_zscore_window_excludes_the_current_index(po):
    # at index 11 the window is [7, 10) (three ones, deviation 0), not
    # [8, 11) (which would hold the replaced 1.45): 1.2 is a signal
    x = _ones(i10=10.0, i11=1.2)
    assert po.zscore_peaks(x, filter_threshold=0.0, **Z).tolist() == [10, 11]


def test_zscore_replacement_feeds_later_windows(po):
    # two spikes: 11 is a signal too (its window [7, 10) is flat); 10 becomes
    # 1.45 and 11 becomes 0.05 * 10 + 0.95 * 1.45 = 1.8775, so the window
    # [9, 12) = 1, 1.45, 1.8775 (mean 1.4425, deviation 0.439) makes 1.9 at
    # 13 no signal (0.4575 < 1.900)
    x = _ones(i10=10.0, i11=10.0, i13=1.9)
    assert po.zscore_peaks(x, filter_threshold=0.0, **Z).tolist() == [10, 11]
    # 3.5 at 14 against [10, 13) = 1.45, 1.8775, 1: 2.0575 > 1.900, a signal;
    # against the unreplaced 10, 10, 1 (mean 7, deviation 5.2) it would not be
    y = _ones(i10=10.0, i11=10.0, i14=3.5)
    assert po.zscore_peaks(y, filter_threshold=0.0, **Z).tolist() == [10, 11, 14]


def test_zscore_below_mean_is_replaced_but_no_signal(po):
    # -5 deviates (flat window) but lies below the mean: no signal; it is
    # replaced by 0.05 * -5 + 0.95 = 0.7, which makes 1.0 at 12 (window
    # [8, 11) = 1, 1, 0.7: mean 0.9, deviation 0.173) no signal either
    x = _ones(i10=-5.0)
    assert po.zscore_peaks(x, filter_threshold=-100.0, **Z).tolist() == []
    # the replaced 0.7 is in the window of 12: 1.7 there deviates 0.8 > 0.75
    y = _ones(i10=-5.0, i12=1.7)
    assert po.zscore_peaks(y, filter_threshold=-100.0, **Z).tolist() == [12]


def test_zscore_filter_skips_without_updating(po):
    # every input <= the filter threshold is skipped: no signal at all
    assert po.zscore_peaks(_ones(i10=10.0), filter_threshold=20.0, **Z).tolist() == []
    # 1.6 at 6 is a signal and becomes 1.03; the window of the last processed
    # index 9, [6, 9) = 1.03, 1, 1 (mean 1.01, deviation 0.0173), stays in
    # force across the skipped 0.5 at 10, so 1.05 at 11 (0.04 < 0.075) is no
    # signal; a window moved to [7, 10) (flat) would make it one
    x = _ones(i6=1.6, i10=0.5, i11=1.05)
    assert po.zscore_peaks(x, filter_threshold=0.9, **Z).tolist() == [6]


def test_zscore_short_input(po):
    # n <= lag + 2: nothing (z-score-peak-extraction.cc:57-58)
    assert po.zscore_peaks(np.array([1.0, 5.0, 1.0, 9.0, 1.0]), filter_threshold=0.0,
                           lag_percentile=0.9, lag_max=3).tolist() == []
