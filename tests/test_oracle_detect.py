"""Pins of the detector oracle (CPU): the Layer-1 sigma-clip estimator, the
Layer-2 sigma_k state, box sums and the top-k + merge-radius decode."""
import math

import numpy as np
from scipy.stats import norm

from oracle import detect as od
from oracle import search as osearch


def test_lower_median_known_answer():
    # lower median of [1, 2, 3, 100] is 2 (a mid-point median gives 2.5);
    # no sample leaves the 3-sigma band, so sigma = sqrt(mean((x - 2)^2))
    assert od.lower_median(np.array([3.0, 1.0, 100.0, 2.0])) == 2.0
    s = od.sigma_clipped_std(np.array([1.0, 2.0, 3.0, 100.0]))
    assert abs(s - math.sqrt((1 + 0 + 1 + 98 ** 2) / 4)) < 1e-12
    assert abs(s - math.sqrt((1.5 ** 2 + 0.5 ** 2 * 2 + 97.5 ** 2) / 4)) > 0.2
    # NaNs are ignored, an empty sample gives 0
    assert abs(od.sigma_clipped_std(np.array([1.0, np.nan, 2.0, 3.0, 100.0])) - s) < 1e-12
    assert od.sigma_clipped_std(np.array([np.nan])) == 0.0


def _clip_fixed_point(n_sigma=3.0, n_iter=3):
    """Clipped std of N(0, 1) after each pass: sigma_{i+1}^2 = 1 - 2 a phi(a) / (2 Phi(a) - 1), a = n_sigma sigma_i."""
    s = 1.0
    for _ in range(n_iter):
        a = n_sigma * s
        s = math.sqrt(1.0 - 2.0 * a * norm.pdf(a) / (2.0 * norm.cdf(a) - 1.0))
    return s


def test_sigma_clip_truncated_normal_fixed_point():
    x = np.random.default_rng(3).normal(0.0, 2.5, 1_000_000)
    got = od.sigma_clipped_std(x)
    want = 2.5 * _clip_fixed_point()
    assert abs(got - want) < 3e-3 * 2.5, (got, want)
    # one pass fewer lands elsewhere (the pin resolves the iteration count)
    assert abs(od.sigma_clipped_std(x, n_iterations=1) - 2.5 * _clip_fixed_point(n_iter=1)) < 3e-3 * 2.5
    assert abs(_clip_fixed_point(n_iter=1) - _clip_fixed_point()) > 1e-3


def test_layer2_burnin_mean_then_ema():
    L = od.Layer2(2, cadence_s=0.5, tau_s=30.0, n_burnin=3, seed=[5.0, 7.0])
    got = [L.update(np.array([x, 2 * x])) for x in (1.0, 2.0, 3.0)]
    # burn-in: running mean of the estimates (the seed carries no weight)
    assert np.allclose(got[0], [1, 2]) and np.allclose(got[1], [1.5, 3]) and np.allclose(got[2], [2, 4])
    assert not L.is_warming_up
    g = 1.0 - math.exp(-0.5 / 30.0)
    s = L.update(np.array([10.0, 0.0]))          # EMA; a zero estimate keeps the old value
    assert np.allclose(s, [g * 10 + (1 - g) * 2, 4.0])
    assert np.allclose(L.update(np.array([10.0, 4.0]), valid=False), s)   # invalid cube: unchanged
    # the EMA converges to a constant input
    for _ in range(20000):
        s = L.update(np.array([3.0, 3.0]))
    assert np.allclose(s, 3.0, rtol=1e-9)


def test_layer2_ceiling_and_escape():
    L = od.Layer2(1, cadence_s=1.0, tau_s=1e9, n_burnin=0, seed=[1.0], sigma_max_ratio=2.0,
                  clamp_escape_cubes=2)
    L.update(np.array([1.0]))
    s1 = L.update(np.array([10.0]))     # clamped to s_k: streak 1
    s2 = L.update(np.array([10.0]))     # clamped: streak 2
    s3 = L.update(np.array([10.0]))     # escape: s_k jumps to the estimate, then the EMA step
    assert np.allclose(s1, 1.0) and np.allclose(s2, 1.0) and np.allclose(s3, 10.0)


def test_box_sums_sliding_windows():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((40, 3))
    B = od.box_sums(x, [1, 2, 4, 8])
    for k, w in enumerate([1, 2, 4, 8]):
        want = osearch.boxcar_np(x, w, axis=0)
        np.testing.assert_allclose(B[k], want, rtol=0, atol=1e-12)


def test_decode_planted_peaks_and_radius_invariants():
    rng = np.random.default_rng(2)
    n_out, F, H, W = 60, 8, 12, 12
    snr = rng.standard_normal((n_out, F, H, W))
    win = rng.integers(0, 4, snr.shape)
    peaks = [(5, 1, 2, 3, 30.0), (40, 6, 9, 9, 25.0), (20, 3, 6, 1, 20.0)]
    for t, f, l, m, v in peaks:
        snr[t, f, l, m] = v
        snr[t + 1, f, l, m] = v - 1        # inside the merge radius of the peak
    got = od.decode_topk(snr, win, threshold=3.0, n_top=64, radius_t=4, radius_fdm=1, radius_lm=1)
    keys = [(t, f, l, m) for t, f, l, m, _, _ in got]
    for t, f, l, m, v in peaks:
        assert (t, f, l, m) in keys and (t + 1, f, l, m) not in keys
    for i, a in enumerate(got):                   # accepted candidates are pairwise outside the radii
        for b in got[:i]:
            assert not (abs(a[0] - b[0]) <= 4 and abs(a[1] - b[1]) <= 1 and abs(a[2] - b[2]) <= 1
                        and abs(a[3] - b[3]) <= 1)
    assert all(v > 3.0 for *_, v in got)
    assert [v for *_, v in got] == sorted([v for *_, v in got], reverse=True)
