"""Comprehensive window features (the reference's comprehensive set,
src/pybasin/ts_torch/settings.py:223-340) on the CPU.

* the oracle's restatements (oracle.comprehensive) are pinned against
  library routines that compute the same quantities another way
  (scipy.stats skew / kurtosis with bias=False, scipy.stats.linregress,
  numpy.correlate, scipy.signal.argrelmax) and against hand-computed
  answers on short series a plausible slip (a strict comparison, an
  off-by-one position, a wrong normaliser) would fail;
* the product's CPU path (the torch restatement of the window kernel's
  slots, then the shared derivation) equals the oracle on random, tied,
  constant and short windows, and gives NaN for windows with a non-finite
  value;
* the extractor names the columns as the reference does and rejects
  parameter values the kernel does not reduce.
"""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pybasin_amd.features import StatFeatureExtractor, Solution
from pybasin_amd.window_features import (AUTOCOV_FEATURES, COMPREHENSIVE_FEATURES,
                                         PEAK_FEATURES, SORTED_FEATURES, SPECTRUM_FEATURES,
                                         TRENDS_FEATURES, _block_cpu, autocov_feature,
                                         check_params, comprehensive_feature, feature_family,
                                         peak_feature, sorted_feature, spectrum_feature,
                                         trends_feature, window_autocov, window_peak_counts,
                                         window_perm_entropy, window_sorted, window_spectrum,
                                         window_trends)


def _all_params():
    """The calculators formed from the window-feature slots."""
    for name, grid in COMPREHENSIVE_FEATURES.items():
        for p in (grid or [None]):
            if feature_family(name, p) == "block":
                yield name, p


def _spectrum_params(nb=None):
    """The spectrum calculators of the comprehensive set, with the bins
    limited to a spread of indices (and nb, nb + 1: past the last bin)."""
    ks = [0, 1, 2, 3, 5, 8, 13, 31, 64, 99] + ([nb, nb + 1] if nb is not None else [])
    for a in ("real", "imag", "abs", "angle"):
        for k in ks:
            yield "fft_coefficient", {"coeff": k, "attr": a}
    for k in ks:
        yield "spkt_welch_density", {"coeff": k}
    for name in ("fft_aggregated", "fourier_entropy"):
        for p in SPECTRUM_FEATURES[name]:
            yield name, p


def test_oracle_moments_and_trend_match_scipy():
    from scipy import stats

    rng = np.random.default_rng(2)
    for n in (7, 50, 1000):
        x = rng.normal(size=n) ** 3 + 0.3 * np.arange(n) / n
        assert O.comprehensive(x, "skewness") == pytest.approx(stats.skew(x, bias=False), rel=1e-8)
        assert O.comprehensive(x, "kurtosis") == pytest.approx(stats.kurtosis(x, bias=False),
                                                               rel=1e-8)
        lr = stats.linregress(np.arange(n), x)
        for attr, want in (("slope", lr.slope), ("intercept", lr.intercept),
                           ("rvalue", lr.rvalue), ("stderr", lr.stderr)):
            assert O.comprehensive(x, "linear_trend", {"attr": attr}) == pytest.approx(
                want, rel=1e-7, abs=1e-12), attr
        # the reference's pseudo p-value 1 / (1 + |t|), t = r sqrt(n - 2) / sqrt(1 - r^2)
        t = lr.rvalue * math.sqrt(n - 2) / math.sqrt(1 - lr.rvalue ** 2)
        assert O.comprehensive(x, "linear_trend", {"attr": "pvalue"}) == pytest.approx(
            1 / (1 + abs(t)), rel=1e-7)


def test_oracle_autocorrelation_and_peaks_match_numpy_scipy():
    from scipy.signal import argrelmax

    rng = np.random.default_rng(3)
    x = np.sin(np.arange(300) * 0.37) + 0.4 * rng.normal(size=300)
    c = x - x.mean()
    full = np.correlate(c, c, mode="full")
    for lag in range(1, 10):
        assert O.comprehensive(x, "autocorrelation", {"lag": lag}) == pytest.approx(
            full[len(x) - 1 + lag] / full[len(x) - 1], rel=1e-9)
    assert O.comprehensive(x, "autocorrelation", {"lag": 0}) == 1.0
    assert O.comprehensive(x[:5], "autocorrelation", {"lag": 7}) == 0.0
    for n in (1, 3, 5):
        idx = argrelmax(x, order=n)[0]
        idx = idx[(idx >= n) & (idx < len(x) - n)]  # the reference never marks the edges
        assert O.comprehensive(x, "number_peaks", {"n": n}) == len(idx)


def test_oracle_partial_autocorrelation_and_binned_entropy():
    rng = np.random.default_rng(4)
    # AR(1) with phi = 0.6: partial autocorrelation 0.6 at lag 1, 0 beyond
    # (within sampling error 3 / sqrt(n))
    e = rng.normal(size=20000)
    x = np.zeros_like(e)
    for i in range(1, len(e)):
        x[i] = 0.6 * x[i - 1] + e[i]
    assert O.comprehensive(x, "partial_autocorrelation", {"lag": 1}) == pytest.approx(0.6, abs=0.03)
    for lag in range(2, 10):
        assert abs(O.comprehensive(x, "partial_autocorrelation", {"lag": lag})) < 0.03
    assert O.comprehensive(x, "partial_autocorrelation", {"lag": 1}) == pytest.approx(
        O.comprehensive(x, "autocorrelation", {"lag": 1}), rel=1e-12)
    # on well-conditioned series the recursion equals the last coefficient of
    # the order-k Yule-Walker system solved directly
    for seed in range(3):
        y = np.random.default_rng(seed).normal(size=400).cumsum() * 0.1 + rng.normal(size=400)
        c = y - y.mean()
        r = np.array([1.0] + [(c[:-l] * c[l:]).sum() / (c @ c) for l in range(1, 10)])
        for k in range(1, 10):
            R = np.array([[r[abs(i - j)] for j in range(k)] for i in range(k)])
            assert np.linalg.cond(R) < 1e4
            assert O.comprehensive(y, "partial_autocorrelation", {"lag": k}) == pytest.approx(
                np.linalg.solve(R, r[1:k + 1])[-1], rel=1e-7, abs=1e-10)
    assert O.comprehensive(x[:4], "partial_autocorrelation", {"lag": 5}) == 0.0
    # entropies: ten equally filled bins -> ln 10; two -> ln 2; constant -> 0
    assert O.comprehensive(np.arange(100.0), "binned_entropy", {"max_bins": 10}) == pytest.approx(
        math.log(10))
    assert O.comprehensive(np.array([0.0, 1.0] * 50), "binned_entropy", {"max_bins": 10}) == \
        pytest.approx(math.log(2))
    assert O.comprehensive(np.full(9, 3.0), "binned_entropy", {"max_bins": 10}) == 0.0
    # the bin of the maximum is the last one, (x - min) / range * (10 - 1e-6) < 10
    assert O.comprehensive(np.array([0.0, 1.0]), "binned_entropy", {"max_bins": 10}) == \
        pytest.approx(math.log(2))


def test_oracle_known_answers():
    C = O.comprehensive
    x = np.array([1.0, 3.0, 3.0, -2.0, 0.0, 3.0, -2.0, 1.0])  # n = 8, mean 7/8
    assert C(x, "sum_values") == 7.0 and C(x, "length") == 8.0
    assert C(x, "abs_energy") == 1 + 9 + 9 + 4 + 0 + 9 + 4 + 1
    assert C(x, "absolute_sum_of_changes") == 2 + 0 + 5 + 2 + 3 + 5 + 3
    assert C(x, "mean_abs_change") == pytest.approx(20 / 7)
    assert C(x, "mean_change") == pytest.approx(0.0)
    assert C(x, "mean_second_derivative_central") == pytest.approx((1 - (-2) - 3 + 1) / 12)
    assert C(x, "count_above_mean") == 5.0 and C(x, "count_below_mean") == 3.0
    assert C(x, "count_above", {"t": 0}) == 5 / 8 and C(x, "count_below", {"t": 0}) == 2 / 8
    assert C(x, "count_value", {"value": 0}) == 1.0 and C(x, "count_value", {"value": 1}) == 2.0
    assert C(x, "count_in_range", {"min_val": -1, "max_val": 1}) == 3.0
    assert C(x, "has_duplicate_max") == 1.0 and C(x, "has_duplicate_min") == 1.0
    assert C(x, "first_location_of_maximum") == 1 / 8 and C(x, "last_location_of_maximum") == 5 / 8
    assert C(x, "first_location_of_minimum") == 3 / 8 and C(x, "last_location_of_minimum") == 6 / 8
    # above the mean 7/8: T T T F F T F T; below: F F F T T F T F
    assert C(x, "longest_strike_above_mean") == 3.0 and C(x, "longest_strike_below_mean") == 2.0
    # above 0: T T T F F T F T -> 4 changes; above 1: F T T F F T F F -> 4; above -1: T T T F T T F T -> 4
    assert C(x, "number_crossing_m", {"m": 0}) == 4.0 and C(x, "number_crossing_m", {"m": 1}) == 4.0
    assert C(x, "number_crossing_m", {"m": -1}) == 4.0
    # c3 lag 1: mean of x0 x1 x2 ... over 6 triples
    trip = [x[i] * x[i + 1] * x[i + 2] for i in range(6)]
    assert C(x, "c3", {"lag": 1}) == pytest.approx(np.mean(trip))
    tr = [x[i + 2] ** 2 * x[i + 1] - x[i + 1] * x[i] ** 2 for i in range(6)]
    assert C(x, "time_reversal_asymmetry_statistic", {"lag": 1}) == pytest.approx(np.mean(tr))
    assert C(x, "c3", {"lag": 4}) == 0.0
    assert C(x, "mean_n_absolute_max", {"number_of_maxima": 3}) == 3.0
    assert C(x, "mean_n_absolute_max", {"number_of_maxima": 5}) == pytest.approx(13 / 5)
    # |x| mass 15: cumulative 1 4 7 9 9 12 14 15 -> q = 0.3 (>= 4.5): index 2; q = 0.6 (>= 9): 3
    assert C(x, "index_mass_quantile", {"q": 0.6}) == 3 / 8
    assert C(x, "index_mass_quantile", {"q": 0.3}) == 2 / 8
    # ramp: every change 1; constant: equal chunks
    r = np.arange(20, dtype=float)
    assert C(r, "cid_ce", {"normalize": False}) == pytest.approx(math.sqrt(19))
    assert C(r, "cid_ce", {"normalize": True}) == pytest.approx(math.sqrt(19) / r.std())
    k = np.full(30, 2.0)
    for j in range(10):
        assert C(k, "energy_ratio_by_chunks", {"num_segments": 10, "segment_focus": j}) == \
            pytest.approx(0.1)
    assert C(k[:9], "energy_ratio_by_chunks", {"num_segments": 10, "segment_focus": 0}) == 0.0
    # ratio beyond r sigma: values 0 x 8, 10 x 2 -> mean 2, sd 4; |x - 2| = 2 (8x), 8 (2x)
    b = np.array([0.0] * 8 + [10.0] * 2)
    assert C(b, "ratio_beyond_r_sigma", {"r": 1}) == 0.2 and C(b, "ratio_beyond_r_sigma", {"r": 0.25}) == 1.0
    assert C(b, "ratio_beyond_r_sigma", {"r": 0.5}) == 0.2  # |x - mean| = r sd exactly is not beyond
    assert C(b, "has_variance_larger_than_standard_deviation") == 1.0
    assert C(b, "large_standard_deviation", {"r": 0.35}) == 1.0
    assert C(b, "large_standard_deviation", {"r": 0.45}) == 0.0
    s = np.array([-1.0, 0.0, 1.0])
    assert C(s, "symmetry_looking", {"r": 0.05}) == 1.0
    assert C(np.array([0.0, 0.0, 3.0]), "symmetry_looking", {"r": 0.25}) == 0.0
    assert math.isnan(C(np.array([1.0, np.inf, 2.0]), "sum_values"))


def test_oracle_spectrum_matches_numpy_rfft():
    """The oracle's explicit DFT sum against numpy's FFT (another algorithm),
    odd and even lengths, every attr, bins past the last one."""
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 64, 255, 1000):
        x = rng.normal(size=n) + 0.3
        X = np.fft.rfft(x)
        nb = X.size
        for k in list(range(min(nb, 12))) + [nb - 1, nb, nb + 7]:
            want = X[k] if k < nb else 0j
            for attr, v in (("real", want.real), ("imag", want.imag), ("abs", abs(want))):
                assert O.comprehensive(x, "fft_coefficient", {"coeff": k, "attr": attr}) == \
                    pytest.approx(v, abs=1e-9 * n), (n, k, attr)
            if abs(want) > 1e-6:
                d = abs(O.comprehensive(x, "fft_coefficient", {"coeff": k, "attr": "angle"})
                        - np.angle(want))
                assert min(d, 2 * np.pi - d) <= 1e-9, (n, k)
            assert O.comprehensive(x, "spkt_welch_density", {"coeff": k}) == pytest.approx(
                abs(want) ** 2 / n, rel=1e-9, abs=1e-9), (n, k)


def test_oracle_spectrum_closed_forms():
    """Pure tones put their power in known bins: one tone -> centroid k,
    variance 0; two tones with powers w1 : w2 -> the two-point law's moments
    (centroid, variance w1 w2 d^2, skew (w1 - w2) / sqrt(w1 w2), kurtosis
    1 / (w1 w2) - 3, entropy -sum w ln w); sine -> angle -pi / 2; DC and
    Nyquist bins of constant and alternating series."""
    C = O.comprehensive
    n = 240
    i = np.arange(n)
    x = np.cos(2 * np.pi * 17 * i / n)
    assert C(x, "fft_coefficient", {"coeff": 17, "attr": "real"}) == pytest.approx(n / 2)
    assert C(x, "fft_aggregated", {"aggtype": "centroid"}) == pytest.approx(17.0)
    assert C(x, "fft_aggregated", {"aggtype": "variance"}) == pytest.approx(0.0, abs=1e-9)
    nb = n // 2 + 1  # one bin holds all power; the other nb - 1 clamp to 1e-10
    assert C(x, "fourier_entropy", {"bins": 10}) == pytest.approx(
        (nb - 1) * 1e-10 * math.log(1e10), rel=1e-3)
    s = np.sin(2 * np.pi * 9 * i / n)
    assert C(s, "fft_coefficient", {"coeff": 9, "attr": "imag"}) == pytest.approx(-n / 2)
    assert C(s, "fft_coefficient", {"coeff": 9, "attr": "angle"}) == pytest.approx(-np.pi / 2)
    a, b, k1, k2 = 1.0, 2.0, 10, 40
    two = a * np.cos(2 * np.pi * k1 * i / n) + b * np.cos(2 * np.pi * k2 * i / n)
    w1, w2 = a * a / (a * a + b * b), b * b / (a * a + b * b)
    d = k2 - k1
    assert C(two, "fft_aggregated", {"aggtype": "centroid"}) == pytest.approx(w1 * k1 + w2 * k2)
    assert C(two, "fft_aggregated", {"aggtype": "variance"}) == pytest.approx(w1 * w2 * d * d)
    assert C(two, "fft_aggregated", {"aggtype": "skew"}) == pytest.approx(
        (w1 - w2) / math.sqrt(w1 * w2), rel=1e-9)
    assert C(two, "fft_aggregated", {"aggtype": "kurtosis"}) == pytest.approx(
        1 / (w1 * w2) - 3, rel=1e-9)
    assert C(two, "fourier_entropy", {"bins": 3}) == pytest.approx(
        -(w1 * math.log(w1) + w2 * math.log(w2)), rel=1e-6)
    assert C(np.full(30, -2.0), "fft_coefficient", {"coeff": 0, "attr": "real"}) == \
        pytest.approx(-60.0)
    assert C(np.full(30, -2.0), "fft_coefficient", {"coeff": 0, "attr": "angle"}) == \
        pytest.approx(np.pi)
    alt = (-1.0) ** np.arange(16)
    assert C(alt, "fft_coefficient", {"coeff": 8, "attr": "real"}) == pytest.approx(16.0)
    assert C(alt, "spkt_welch_density", {"coeff": 8}) == pytest.approx(16.0)
    assert C(alt, "fft_coefficient", {"coeff": 9, "attr": "abs"}) == 0.0


def _spectrum_close(g, want, name, p, scale):
    """The spectrum tolerance: bins to 1e-9 of the series' sum |x| (angles
    only where the bin is well above that), the normalised aggregates to
    1e-9."""
    if not math.isfinite(want):
        return not math.isfinite(g)
    if name == "fft_coefficient":
        if p["attr"] == "angle":
            return True  # checked through real and imag
        return abs(g - want) <= 1e-9 * scale
    if name == "spkt_welch_density":
        return abs(g - want) <= 1e-9 * scale * scale
    return g == pytest.approx(want, rel=1e-9, abs=1e-9)


def _angle_ok(got_re, got_im, x, k, scale):
    X = O.comprehensive(x, "fft_coefficient", {"coeff": k, "attr": "abs"})
    if X <= 1e-6 * scale:
        return True
    want = O.comprehensive(x, "fft_coefficient", {"coeff": k, "attr": "angle"})
    d = abs(math.atan2(got_im, got_re) - want)
    return min(d, 2 * math.pi - d) <= 1e-6


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_spectrum_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        nb = N // 2 + 1
        plist = list(_spectrum_params(nb))
        coeff, agg = window_spectrum(wt, 0, 102 + nb, True)
        for name, p in plist:
            got = spectrum_feature(name, p, coeff, agg, N).numpy()
            for b in range(B):
                for s in range(S):
                    x = x64[:, b, s]
                    want = O.comprehensive(x, name, p)
                    scale = max(np.abs(x).sum(), 1e-300)
                    assert _spectrum_close(float(got[b, s]), want, name, p, scale), \
                        (name, p, N, b, s, float(got[b, s]), want)
                    if name == "fft_coefficient" and p["attr"] == "angle":
                        k = p["coeff"]
                        assert _angle_ok(float(coeff[min(k, coeff.shape[0] - 1), 0, b, s]),
                                         float(coeff[min(k, coeff.shape[0] - 1), 1, b, s]),
                                         x, k, scale), (N, k, b, s)


def test_oracle_sorted_family_known_answers():
    """Hand-counted multiplicities and a change-quantile corridor whose
    bounds equal sample values (a strict comparison would drop them)."""
    C = O.comprehensive
    x = np.array([1.0, 2, 2, 3, 3, 3, 5])  # distinct 1, 2, 3, 5; 2 and 3 reoccur
    assert C(x, "has_duplicate") == 1.0
    assert C(x, "ratio_value_number_to_time_series_length") == pytest.approx(4 / 7)
    assert C(x, "percentage_of_reoccurring_datapoints_to_all_datapoints") == pytest.approx(2 / 4)
    assert C(x, "percentage_of_reoccurring_values_to_all_values") == pytest.approx(5 / 7)
    assert C(x, "sum_of_reoccurring_data_points") == 13.0
    assert C(x, "sum_of_reoccurring_values") == 5.0
    # quantiles 0.2 and 0.8 of 7 values: ranks 1.2 and 4.8 -> 2 and 3; the
    # samples 2, 2, 3, 3, 3 lie in [2, 3]; their changes 0, 1, 0, 0
    cq = lambda isabs, f: C(x, "change_quantiles", {"ql": 0.2, "qh": 0.8, "isabs": isabs,  # noqa: E731
                                                     "f_agg": f})
    assert cq(True, "mean") == pytest.approx(0.25)
    assert cq(True, "var") == pytest.approx(0.1875)
    y = np.array([0.0, 4.0, 1.0, 3.0, 2.0])  # changes 4, -3, 2, -1
    # ql = 0 and qh = 1 take every change
    assert C(y, "change_quantiles", {"ql": 0.0, "qh": 1.0, "isabs": False, "f_agg": "mean"}) == \
        pytest.approx(0.5)
    assert C(y, "change_quantiles", {"ql": 0.0, "qh": 1.0, "isabs": True, "f_agg": "mean"}) == \
        pytest.approx(2.5)
    # quantile 0.6 of 5 values: rank 2.4 -> 2.4; each adjacent pair (0, 4), (4, 1),
    # (1, 3), (3, 2) holds a sample above it, so no change counts
    assert C(y, "change_quantiles", {"ql": 0.0, "qh": 0.6, "isabs": True, "f_agg": "mean"}) == 0.0
    d = np.arange(12.0)
    assert C(d, "has_duplicate") == 0.0 and C(d, "sum_of_reoccurring_values") == 0.0
    assert C(np.full(9, 2.5), "percentage_of_reoccurring_values_to_all_values") == 1.0
    assert C(np.full(9, 2.5), "sum_of_reoccurring_data_points") == 22.5
    assert C(np.array([3.0]), "percentage_of_reoccurring_values_to_all_values") == 0.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_quantile_matches_torch(dtype):
    """The oracle's quantile (the corridor bounds of change_quantiles)
    equals torch.quantile in the series dtype, ties and odd sizes included."""
    rng = np.random.default_rng(8)
    for n in (1, 2, 5, 7, 100, 999, 1000):
        for x in (rng.normal(size=n), np.round(rng.normal(size=n) * 2) / 2):
            xs = x.astype(dtype)
            srt = np.sort(xs)
            for q in (0.0, 0.1, 0.2, 0.4, 0.6, 0.8, 0.9, 1.0):
                want = torch.quantile(torch.from_numpy(xs), q).item()
                got = O._torch_quantile(srt, q, dtype)
                assert got == pytest.approx(want, rel=4 * np.finfo(dtype).eps, abs=0), (n, q)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_sorted_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        dup, cq, _ = window_sorted(wt, 0, True)
        for name, grid in SORTED_FEATURES.items():
            for p in (grid or [None]):
                got = sorted_feature(name, p, dup, cq, N).numpy()
                for b in range(B):
                    for s in range(S):
                        want = O.comprehensive(wt[:, b, s].numpy(), name, p,
                                               np.float32 if dtype == torch.float32 else np.float64)
                        assert got[b, s] == pytest.approx(want, rel=1e-9, abs=1e-12), \
                            (name, p, N, b, s)


def test_oracle_autocov_family_pins():
    """agg_autocorrelation against the reference's own route (the inverse FFT
    of the zero-padded power spectrum, by numpy), maxlag past the window;
    ar_coefficient against scipy's Levinson Toeplitz solver and an AR(2)
    process's known coefficients; augmented_dickey_fuller against
    scipy.stats.linregress of the changes on the lagged values."""
    from scipy import linalg, stats

    rng = np.random.default_rng(12)
    for n in (20, 300):
        x = np.cumsum(rng.normal(size=n)) * 0.1 + np.sin(np.arange(n) * 0.3)
        xc = x - x.mean()
        C = np.fft.irfft(np.abs(np.fft.rfft(xc, 2 * n)) ** 2, 2 * n)[:41]
        L = min(40, n - 1)
        acf = C[1:L + 1] / (C[0] + 1e-10)
        for f, want in (("mean", acf.mean()), ("median", np.sort(acf)[(L - 1) // 2]),
                        ("var", acf.var())):
            assert O.comprehensive(x, "agg_autocorrelation", {"f_agg": f, "maxlag": 40}) == \
                pytest.approx(want, rel=1e-9, abs=1e-12), (n, f)
        a = C[:11] / (C[0] + 1e-10)
        if n < 11:
            a[n:] = 0.0
        col = a[:10].copy()
        col[0] += 1e-6
        phi = linalg.solve_toeplitz(col, a[1:11])
        for c in range(10):
            assert O.comprehensive(x, "ar_coefficient", {"coeff": c, "k": 10}) == pytest.approx(
                phi[c], rel=1e-7, abs=1e-10), (n, c)
        assert O.comprehensive(x, "ar_coefficient", {"coeff": 10, "k": 10}) == 0.0
        lr = stats.linregress(x[:-1], np.diff(x))
        t = O.comprehensive(x, "augmented_dickey_fuller", {"attr": "teststat"})
        assert t == pytest.approx(lr.slope / lr.stderr, rel=1e-7)
        assert O.comprehensive(x, "augmented_dickey_fuller", {"attr": "pvalue"}) == \
            pytest.approx(1 / (1 + abs(t)))
        assert O.comprehensive(x, "augmented_dickey_fuller", {"attr": "usedlag"}) == 1.0
    e = rng.normal(size=40_000)
    ar = np.zeros_like(e)
    for i in range(2, e.size):
        ar[i] = 0.6 * ar[i - 1] - 0.3 * ar[i - 2] + e[i]
    got = [O.comprehensive(ar, "ar_coefficient", {"coeff": c, "k": 10}) for c in range(10)]
    assert got[0] == pytest.approx(0.6, abs=0.03) and got[1] == pytest.approx(-0.3, abs=0.03)
    assert max(abs(v) for v in got[2:]) < 0.03


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_autocov_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        C = window_autocov(wt, 0, 40)
        memo: dict = {}
        for name, grid in AUTOCOV_FEATURES.items():
            for p in grid:
                got = autocov_feature(name, p, C, N, memo).numpy()
                for b in range(B):
                    for s in range(S):
                        want = O.comprehensive(x64[:, b, s], name, p)
                        if not math.isfinite(want):
                            assert not math.isfinite(got[b, s]), (name, p, N)
                            continue
                        # the Durbin-Levinson recursion against the oracle's LU solve
                        tol = 1e-7 if name == "ar_coefficient" else 1e-12
                        assert got[b, s] == pytest.approx(want, rel=1e-7, abs=tol), \
                            (name, p, N, b, s)


def test_oracle_trends_family_pins():
    """agg_linear_trend against scipy.stats.linregress of numpy's chunk
    aggregates; the cubic fit against numpy.polyfit (the 1e-6 ridge is
    negligible here) and its sign-change scan against numpy's sign changes;
    benford_correlation against 8/9 of numpy.corrcoef over digits read from
    the decimal strings; cwt_coefficients against numpy.convolve of the
    reflect-padded series with the Ricker wavelet, and a constant series'
    coefficients against the wavelet's sum."""
    from scipy import stats

    rng = np.random.default_rng(13)
    x = np.cumsum(rng.normal(size=503)) * 0.2 + 3 * np.sin(np.arange(503) * 0.05)
    for cs in (5, 10, 50):
        nc = 503 // cs
        ch = x[: nc * cs].reshape(nc, cs)
        for f, a in (("max", ch.max(1)), ("min", ch.min(1)), ("mean", ch.mean(1)), ("var", ch.var(1))):
            lr = stats.linregress(np.arange(nc), a)
            for attr, want in (("slope", lr.slope), ("intercept", lr.intercept),
                               ("rvalue", lr.rvalue), ("stderr", lr.stderr)):
                got = O.comprehensive(x, "agg_linear_trend", {"attr": attr, "chunk_size": cs, "f_agg": f})
                assert got == pytest.approx(want, rel=1e-7, abs=1e-12), (cs, f, attr)
    assert O.comprehensive(x[:9], "agg_linear_trend", {"attr": "slope", "chunk_size": 5,
                                                       "f_agg": "max"}) == 0.0  # one chunk
    pos, vel = x[:-1], np.diff(x)
    pf = np.polyfit(pos, vel, 3)
    for c in range(4):
        assert O.comprehensive(x, "friedrich_coefficients", {"coeff": c, "m": 3, "r": 30}) == \
            pytest.approx(pf[c], rel=1e-4, abs=1e-7), c
    cf = [O.comprehensive(x, "friedrich_coefficients", {"coeff": c, "m": 3, "r": 30}) for c in range(4)]
    fit = np.polyval(cf, pos)
    sgn = np.nonzero(np.sign(fit[:-1]) * np.sign(fit[1:]) < 0)[0]
    want = pos[sgn].max() if sgn.size else 0.0
    assert O.comprehensive(x, "max_langevin_fixed_point", {"m": 3, "r": 30}) == pytest.approx(want)
    y = rng.lognormal(0, 2, size=400) * rng.choice([-1, 1], size=400)
    y[::7] = 0.0
    nz = y[y != 0]
    dg = np.array([int(str(f"{abs(v):e}")[0]) for v in nz])
    fr = np.array([(dg == d).mean() for d in range(1, 10)])
    bf = np.log10(1 + 1 / np.arange(1, 10))
    assert O.comprehensive(y, "benford_correlation") == pytest.approx(
        8 / 9 * np.corrcoef(fr, bf)[0, 1], rel=1e-6)
    assert O.comprehensive(np.r_[np.zeros(20), np.ones(9)], "benford_correlation") == 0.0
    # the reference's digit floor(x / 10^floor(log10 x)): just below a power of
    # ten, log10 rounds up to the integer and the digit is 0, not counted
    assert O.comprehensive(np.full(12, 999999.9999999999), "benford_correlation") == 0.0
    assert O.comprehensive(np.r_[np.full(12, 999999.9999999999), np.full(10, 2.0)],
                           "benford_correlation") == O.comprehensive(np.full(10, 2.0),
                                                                     "benford_correlation")
    for wd in (2, 5, 10, 20):
        t = np.arange(-4 * wd, 4 * wd + 1)
        psi = 2 / (np.sqrt(3 * wd) * np.pi ** 0.25) * (1 - (t / wd) ** 2) * np.exp(-t ** 2 / (2 * wd ** 2))
        conv = np.convolve(np.pad(x, 4 * wd, mode="reflect"), psi, mode="valid")
        for c in (0, 3, 14):
            assert O.comprehensive(x, "cwt_coefficients", {"widths": (wd,), "coeff": c, "w": wd}) == \
                pytest.approx(conv[c], rel=1e-9, abs=1e-12), (wd, c)
        assert O.comprehensive(np.full(60, 1.5), "cwt_coefficients",
                               {"widths": (wd,), "coeff": 4, "w": wd}) == pytest.approx(1.5 * psi.sum())
    assert O.comprehensive(x, "cwt_coefficients", {"widths": (5,), "coeff": 1, "w": 2}) == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_trends_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        out = window_trends(wt, 0, True)
        for name, grid in TRENDS_FEATURES.items():
            for p in (grid or [None]):
                got = trends_feature(name, p, out).numpy()
                for b in range(B):
                    for s in range(S):
                        want = O.comprehensive(x64[:, b, s], name, p)
                        g = float(got[b, s])
                        if not math.isfinite(want):
                            assert not math.isfinite(g), (name, p, N)
                            continue
                        tol = 1e-7 if name in ("friedrich_coefficients",) else 1e-11
                        if name == "agg_linear_trend" and p["attr"] == "stderr":
                            tol = 1e-7  # a perfect fit's residual is rounding, under a square root
                        assert g == pytest.approx(want, rel=1e-7, abs=tol), (name, p, N, b, s, g, want)


def test_oracle_peaks_and_patterns_pins():
    """number_peaks of wide support against scipy.signal.argrelmax (edges
    left out as the reference does); number_cwt_peaks against
    scipy.ndimage.maximum_filter1d; permutation_entropy's pattern ids on
    hand-built series: a monotone series has entropy 0, and the windows
    (0, 1, 3) and (1, 3, 2) share the reference's id 3 (a Lehmer code would
    tell them apart and give ln 2 / ln 6)."""
    from scipy.ndimage import maximum_filter1d
    from scipy.signal import argrelmax

    rng = np.random.default_rng(14)
    x = np.sin(np.arange(800) * 0.07) + 0.3 * rng.normal(size=800)
    for k in (10, 50):
        idx = argrelmax(x, order=k)[0]
        want = np.sum((idx >= k) & (idx < x.size - k))
        assert O.comprehensive(x, "number_peaks", {"n": k}) == want
    xr = np.round(x * 4) / 4  # plateaus: equal maxima count for the cwt peaks
    for mw in (1, 5, 8):
        tot = 0
        for wd in range(1, mw + 1):
            mf = maximum_filter1d(xr, size=2 * wd + 1, mode="nearest")
            tot += np.sum(xr[wd:-wd] == mf[wd:-wd])
        assert O.comprehensive(xr, "number_cwt_peaks", {"max_width": mw}) == pytest.approx(tot / mw)
    assert O.comprehensive(x[:3], "number_cwt_peaks", {"max_width": 5}) == \
        pytest.approx(float(x[1] == x[:3].max()))  # only w = 1 fits
    for d in range(3, 8):
        assert O.comprehensive(np.arange(30.0), "permutation_entropy", {"tau": 1, "dimension": d}) == 0.0
    assert O.comprehensive(np.array([0.0, 1, 3, 2]), "permutation_entropy", {"tau": 1, "dimension": 3}) == 0.0
    assert O.comprehensive(np.array([0.0, 2, 1, 3]), "permutation_entropy", {"tau": 1, "dimension": 3}) == \
        pytest.approx(math.log(2) / math.log(6))
    assert O.comprehensive(np.array([0.0, 1.0]), "permutation_entropy", {"tau": 1, "dimension": 3}) == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_peaks_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        c = window_peak_counts(wt, 0, [10, 50])
        counts = {10: c[0], 50: c[1], "widths": c[4:]}
        perm = window_perm_entropy(wt, 0)
        plist = [("number_peaks", {"n": 10}), ("number_peaks", {"n": 50})] + \
            [(nm, p) for nm, g in PEAK_FEATURES.items() for p in g]
        for name, p in plist:
            got = peak_feature(name, p, counts, perm, N).numpy()
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(x64[:, b, s], name, p)
                    assert got[b, s] == pytest.approx(want, rel=1e-12, abs=1e-12), (name, p, N, b, s)


def test_oracle_embedding_pins():
    """correlation_dimension: a straight line and the logistic map's orbit
    (one-dimensional sets) give about 1, white noise more; its correlation
    sums equal scipy's pdist counts (ordered pairs: twice the unordered).
    lyapunov_r: a sinusoid's divergence slope is about 0, the chaotic
    logistic map's clearly positive."""
    from scipy.spatial.distance import pdist

    C = O.comprehensive
    logistic = np.empty(400)
    logistic[0] = 0.3
    for i in range(1, 400):
        logistic[i] = 4 * logistic[i - 1] * (1 - logistic[i - 1])
    assert C(np.arange(300.0), "correlation_dimension", {}) == pytest.approx(1.0, abs=0.1)
    assert C(logistic, "correlation_dimension", {}) == pytest.approx(1.0, abs=0.15)
    noise = np.random.default_rng(5).normal(size=400)
    assert C(noise, "correlation_dimension", {}) > 2.0
    x = noise[:200]
    m = x.size - 3
    orb = np.stack([x[e: e + m] for e in range(4)], 1)
    dd = pdist(orb)
    sd = x.std(ddof=1)
    L0, L1 = np.log(0.1 * sd + 1e-10), np.log(0.5 * sd + 1e-10)
    lr = L0 + np.arange(50) / 49 * (L1 - L0)
    c = np.array([2 * (dd <= np.exp(v)).sum() for v in lr]) / (m * (m - 1))
    ok = c > 0
    want = np.polyfit(lr[ok], np.log(c[ok] + 1e-30), 1)[0]
    assert C(x, "correlation_dimension", {}) == pytest.approx(want, rel=1e-9)
    assert abs(C(np.sin(np.arange(600) * 0.2), "lyapunov_r", {"tau": 1.0})) < 0.01
    assert C(logistic, "lyapunov_r", {"tau": 1.0}) > 0.05
    assert C(logistic, "lyapunov_r", {"tau": 2.0}) == pytest.approx(
        C(logistic, "lyapunov_r", {"tau": 1.0}) / 2)
    assert math.isnan(C(np.arange(8.0), "lyapunov_r", {"tau": 1.0}))  # no embedded vector


def test_oracle_lyapunov_e_pins():
    """Eckmann's exponents: the largest is clearly positive for the logistic
    map (ln 2 per step), all are at most about 0 for a sinusoid; a
    non-finite series gives inf, as the reference; tau divides."""
    logistic = np.empty(500)
    logistic[0] = 0.3
    for i in range(1, 500):
        logistic[i] = 4 * logistic[i - 1] * (1 - logistic[i - 1])
    ex = O.lyapunov_e(logistic)
    assert max(ex) > 0.2
    sine = np.sin(np.arange(400) * 0.3) + 1e-3 * np.random.default_rng(2).normal(size=400)
    assert max(O.lyapunov_e(sine)) < 0.1
    assert O.lyapunov_e(np.r_[logistic[:50], np.nan]) == [math.inf] * 4
    assert O.lyapunov_e(logistic, tau=2.0) == pytest.approx([v / 2 for v in ex])
    assert O.lyapunov_e(np.arange(12.0)) == [math.inf] * 4  # no orbit point


def _rosenstein_by_scipy(x):
    """Rosenstein's slope from other library routines: the mean frequency
    from scipy's periodogram of the (2n - 1)-point zero-padded series (the
    one-sided doubling cancels in the ratio: the length is odd), the nearest
    neighbours from a k-d tree queried for enough candidates to clear the
    temporal band, the divergence slope by numpy.polyfit."""
    from scipy.signal import periodogram
    from scipy.spatial import cKDTree

    n = x.size
    f, P = periodogram(x, nfft=2 * n - 1, detrend=False, window="boxcar")
    mf = (f[1:] * P[1:]).sum() / P[1:].sum()
    tsep = min(max(int(np.ceil(1.0 / mf)), 1), int(0.25 * n))
    m = n - 9
    orb = np.stack([x[e: e + m] for e in range(10)], 1)
    ntraj = m - 19
    tree = cKDTree(orb[:ntraj])
    dist, idx = tree.query(orb[:ntraj], k=min(ntraj, 2 * tsep + 40))
    nb = np.empty(ntraj, dtype=int)
    for i in range(ntraj):
        ok = np.abs(idx[i] - i) > tsep
        nb[i] = idx[i][ok][0]
    div = [np.log(np.sqrt(((orb[np.arange(ntraj) + k] - orb[nb + k]) ** 2).sum(1))).mean()
           for k in range(20)]
    return np.polyfit(np.arange(20.0), div, 1)[0]


def _eckmann_by_scipy(x):
    """Eckmann's exponents from other library routines: neighbours within the
    8th Chebyshev distance from a k-d tree (p = inf), the companion row from
    the reference's ridge normal equations (+ 1e-10 I) by a Cholesky solve,
    the QR steps by scipy.linalg.qr."""
    from scipy.linalg import cho_factor, cho_solve, qr
    from scipy.spatial import cKDTree

    m, no = 3, x.size - 12
    orb = np.stack([x[e: e + no] for e in range(10)], 1)
    tree = cKDTree(orb)
    q = np.eye(4)
    lexp, cnt = np.zeros(4), np.zeros(4)
    for i in range(no):
        d, _ = tree.query(orb[i], k=9, p=np.inf)
        r = d[8]  # the first is the point itself
        nb = [j for j in tree.query_ball_point(orb[i], r, p=np.inf) if j != i]
        X = np.stack([x[np.array(nb) + m * k] - x[i + m * k] for k in range(4)], 1)
        b = x[np.array(nb) + 4 * m] - x[i + 4 * m]
        a = cho_solve(cho_factor(X.T @ X + 1e-10 * np.eye(4)), X.T @ b)
        t = np.zeros((4, 4))
        t[:3, 1:] = np.eye(3)
        t[3] = a
        qq, rr = qr(t @ q)
        sg = np.where(np.sign(np.diag(rr)) == 0, 1.0, np.sign(np.diag(rr)))
        q = qq * sg
        dg = np.diag(rr) * sg
        lexp[dg > 0] += np.log(dg[dg > 0])
        cnt[dg > 0] += 1
    return lexp / cnt / m


def test_oracle_lyapunov_estimates_match_library_routes():
    """The oracle's Rosenstein and Eckmann estimates against the same methods
    put together from scipy's periodogram, k-d trees, a Cholesky solve and QR
    (no loop of the oracle retyped): a wrong temporal band, a neighbour off
    by one, a stepped offset or a missing sign fix would show."""
    rng = np.random.default_rng(16)
    logistic = np.empty(420)
    logistic[0] = 0.3
    for i in range(1, 420):
        logistic[i] = 4 * logistic[i - 1] * (1 - logistic[i - 1])
    noisy_sine = np.sin(np.arange(400) * 0.23) + 0.1 * rng.normal(size=400)
    for x in (logistic, noisy_sine, rng.normal(size=300)):
        assert O.comprehensive(x, "lyapunov_r", {"tau": 1.0}) == pytest.approx(
            _rosenstein_by_scipy(x), rel=1e-8, abs=1e-10)
        np.testing.assert_allclose(O.lyapunov_e(x), _eckmann_by_scipy(x), rtol=1e-6, atol=1e-8)


def test_cpu_embedding_path_equals_oracle():
    from pybasin_amd.window_features import window_embedding

    rng = np.random.default_rng(15)
    w = np.concatenate([np.sin(np.arange(180)[:, None, None] * rng.uniform(0.05, 0.5, size=(1, 3, 1)))
                        + 0.05 * rng.normal(size=(180, 3, 1)), rng.normal(size=(180, 1, 1))], 1)
    wt = torch.tensor(w)
    cd = window_embedding(wt, 0, "corr_dim").numpy()
    lr = window_embedding(wt, 0, "lyap_r").numpy()
    for b in range(4):
        assert cd[b, 0] == pytest.approx(O.comprehensive(w[:, b, 0], "correlation_dimension", {}),
                                         rel=1e-9)
        assert lr[b, 0] == pytest.approx(O.comprehensive(w[:, b, 0], "lyapunov_r", {"tau": 1.0}),
                                         rel=1e-9, abs=1e-12)
    le = window_embedding(wt[:90], 0, "lyap_e").numpy()
    for b in range(4):
        np.testing.assert_allclose(le[:, b, 0], O.lyapunov_e(w[:90, b, 0]), rtol=1e-8, atol=1e-10)


def test_oracle_lempel_ziv_known_answers():
    """Hand-parsed phrases: 0|1|00|11|01 then a seen tail 00 (5 phrases of
    10 symbols); a constant series parses as 0|00|000|0000 (4 of 10 or 12
    symbols: the seen tail 00 of the 12 is not counted); the count is over
    n / ln n."""
    C = O.comprehensive
    x = np.array([0, 1, 0, 0, 1, 1, 0, 1, 0, 0.0])
    assert C(x, "lempel_ziv_complexity", {"bins": 2}) == pytest.approx(5 * math.log(10) / 10)
    assert C(np.zeros(10), "lempel_ziv_complexity", {"bins": 3}) == pytest.approx(4 * math.log(10) / 10)
    assert C(np.zeros(12), "lempel_ziv_complexity", {"bins": 3}) == pytest.approx(4 * math.log(12) / 12)
    # three bins between 0 and 2: symbols 0, 1, 2 (2 clamped from 3 * 2 / 2)
    assert C(np.array([0.0, 1.0, 2.0]), "lempel_ziv_complexity", {"bins": 3}) == \
        pytest.approx(3 * math.log(3) / 3)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_lempel_ziv_path_equals_oracle(dtype):
    from pybasin_amd.window_features import window_lempel_ziv

    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        N, B, S = w.shape
        bins = [2, 3, 5, 10, 100]
        out = window_lempel_ziv(wt, 0, bins).numpy()
        for k, nb in enumerate(bins):
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(wt[:, b, s].numpy(), "lempel_ziv_complexity", {"bins": nb},
                                           np.float32 if dtype == torch.float32 else np.float64)
                    assert out[k, b, s] == pytest.approx(want, rel=1e-12), (nb, N, b, s)


def _windows():
    rng = np.random.default_rng(11)
    out = [rng.normal(size=(200, 3, 2)),
           np.round(rng.normal(size=(120, 2, 2)) * 2) / 2,  # ties and repeated extremes
           np.full((50, 1, 2), 0.75),  # constant
           np.sin(np.arange(997) * 0.05)[:, None, None] * np.ones((1, 2, 1))]
    for n in (1, 2, 3, 5, 9, 11, 12):
        out.append(rng.normal(size=(n, 2, 1)))
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cpu_path_equals_oracle(dtype):
    for w in _windows():
        wt = torch.tensor(w, dtype=dtype)
        x64 = wt.double().numpy()
        N, B, S = w.shape
        blk = _block_cpu(wt)
        med = wt.median(0).values
        for name, p in _all_params():
            if name in ("skewness", "kurtosis", "augmented_dickey_fuller") and N <= 3:
                continue  # the 1e-10 guards of n - 2 and (n - 2)(n - 3), and the
                # division by n - 3, turn rounding noise into the result
            got = comprehensive_feature(name, p, blk, N, med).numpy()
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(x64[:, b, s], name, p,
                                           np.float32 if dtype == torch.float32 else np.float64)
                    g = got[b, s]
                    if not math.isfinite(want):
                        # degenerate windows (a division by n - 2 = 0 and the like): both
                        # sides non-finite (NaN and +-inf are imputed alike)
                        assert not math.isfinite(g), (name, p, N)
                    else:
                        assert g == pytest.approx(want, rel=1e-9, abs=1e-12), (name, p, N, b, s)


def test_cpu_path_nonfinite_window_is_nan():
    w = torch.randn(40, 2, 1, dtype=torch.float64)
    w[7, 1, 0] = float("inf")
    blk = _block_cpu(w)
    assert torch.isnan(blk[:, 1, 0]).all() and torch.isfinite(blk[:, 0, 0]).all()


def test_extractor_names_and_rejects_unreduced_params():
    torch.manual_seed(1)
    t = torch.linspace(0, 10, 200, dtype=torch.float64)
    y = torch.randn(200, 4, 2, dtype=torch.float64)
    sol = Solution(torch.zeros(4, 2), t, y)
    feats = {"kurtosis": None, "autocorrelation": [{"lag": 3}], "linear_trend": [{"attr": "stderr"}],
             "symmetry_looking": [{"r": 0.1}], "mean": None}
    fx = StatFeatureExtractor(time_steady=5.0, features=feats)
    assert fx.needs_window
    f = fx.extract_features(sol)
    assert fx.feature_names[:5] == ["state_0__kurtosis", "state_0__autocorrelation__lag_3",
                                    "state_0__linear_trend__attr_stderr",
                                    "state_0__symmetry_looking__r_0.1", "state_0__mean"]
    start = int((t >= 5.0).nonzero()[0, 0])
    x = y[start:, 2, 1].numpy()
    got = dict(zip(fx.feature_names, f[2].tolist()))
    assert got["state_1__kurtosis"] == pytest.approx(O.comprehensive(x, "kurtosis"), rel=1e-5)
    assert got["state_1__linear_trend__attr_stderr"] == pytest.approx(
        O.comprehensive(x, "linear_trend", {"attr": "stderr"}), rel=1e-5)
    for bad in ({"count_above": [{"t": 0.5}]}, {"autocorrelation": [{"lag": 12}]},
                {"linear_trend": [{"attr": "bogus"}]}, {"fft_coefficient": [{"coeff": -1}]},
                {"spkt_welch_density": [{"coeff": 5000}]}):
        with pytest.raises(ValueError):
            StatFeatureExtractor(features=bad).extract_features(sol)
    for name, p in _all_params():
        check_params(name, p)
    for name, grid in SPECTRUM_FEATURES.items():
        for p in grid:
            check_params(name, p)


def test_extractor_spectrum_columns():
    """The spectrum calculators through the extractor: one rfft per
    extraction, bins past the last one are 0, columns equal the oracle."""
    torch.manual_seed(3)
    t = torch.linspace(0, 10, 101, dtype=torch.float64)
    y = torch.randn(101, 3, 2, dtype=torch.float64)
    sol = Solution(torch.zeros(3, 2), t, y)
    feats = {"fft_coefficient": [{"coeff": 3, "attr": "abs"}, {"coeff": 60, "attr": "real"}],
             "spkt_welch_density": [{"coeff": 2}], "fft_aggregated": [{"aggtype": "skew"}],
             "fourier_entropy": [{"bins": 5}], "mean": None}
    fx = StatFeatureExtractor(time_steady=0.0, features=feats)
    assert fx.needs_window
    f = fx.extract_features(sol)
    assert fx.feature_names[:5] == ["state_0__fft_coefficient__attr_abs__coeff_3",
                                    "state_0__fft_coefficient__attr_real__coeff_60",
                                    "state_0__spkt_welch_density__coeff_2",
                                    "state_0__fft_aggregated__aggtype_skew",
                                    "state_0__fourier_entropy__bins_5"]
    got = dict(zip(fx.feature_names, f[1].tolist()))
    x = y[:, 1, 1].numpy()
    assert got["state_1__fft_coefficient__attr_abs__coeff_3"] == pytest.approx(
        O.comprehensive(x, "fft_coefficient", {"coeff": 3, "attr": "abs"}), rel=1e-6)
    assert got["state_1__fft_coefficient__attr_real__coeff_60"] == 0.0
    assert got["state_1__fft_aggregated__aggtype_skew"] == pytest.approx(
        O.comprehensive(x, "fft_aggregated", {"aggtype": "skew"}), rel=1e-5)
    assert got["state_1__fourier_entropy__bins_5"] == pytest.approx(
        O.comprehensive(x, "fourier_entropy", {"bins": 5}), rel=1e-6)


def test_every_comprehensive_calculator_through_the_extractor():
    """The extractor takes every calculator of the comprehensive grid built
    here on one window (one kernel output per family) and names its columns
    as the reference does."""
    from pybasin_amd.window_features import WindowBlocks

    torch.manual_seed(4)
    t = torch.linspace(0, 30, 300, dtype=torch.float64)
    y = torch.randn(300, 3, 2, dtype=torch.float64).cumsum(0) * 0.1
    sol = Solution(torch.zeros(3, 2), t, y)
    fx = StatFeatureExtractor(time_steady=0.0, features=COMPREHENSIVE_FEATURES)
    f = fx.extract_features(sol)
    n_cols = sum(len(g or [None]) for g in COMPREHENSIVE_FEATURES.values()) + 3  # lyapunov_e: 4
    assert f.shape == (3, 2 * n_cols)
    assert torch.isfinite(f).all()
    assert len(set(fx.feature_names)) == f.shape[1]
    assert "state_1__cwt_coefficients__coeff_3__w_5__widths_(5,)" in fx.feature_names
    assert "state_0__change_quantiles__f_agg_var__isabs_True__qh_0.8__ql_0.2" in fx.feature_names
    assert "state_1__lyapunov_e__emb_dim_10__matrix_dim_4__min_nb_8__min_tsep_0__tau_1.0__3" in \
        fx.feature_names
    wb = WindowBlocks(y, 0, [(0, n, p) for n, g in COMPREHENSIVE_FEATURES.items() for p in (g or [None])])
    assert wb.feature("number_peaks", {"n": 50}).shape == (3, 2)
