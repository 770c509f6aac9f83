"""Comprehensive window features: every calculator of the reference's
comprehensive feature set (src/pybasin/ts_torch/settings.py:223-340) on the
kept window rows [start, N) of a trajectory tensor (N, B, S).

Each family is one kernel of csrc/window.hip (entry points in
include/basinhip.h), run once per extraction when a feature of it is first
asked for (WindowBlocks):
* bh_window_features: every series reduced to raw slots (sums, extremes and
  their positions, changes, counts, crossings, chunk energies, lag products,
  central moments, runs, autocovariances 1..9, index-mass positions, binned
  entropy bins) from which the moment, count, location, trend and
  autocorrelation calculators follow here in float64;
* bh_window_spectrum: rfft bins and the aggregated power spectrum;
* bh_window_sorted + bh_window_change_quantiles: duplicates, reoccurrence
  counts, the median and the change-quantile corridors of the sorted window;
* bh_window_autocov: autocovariances of many lags (agg_autocorrelation,
  ar_coefficient);
* bh_window_trends: chunk trends, the cubic fits, leading digits, wavelets;
* bh_window_peak_counts, bh_window_perm_entropy: wide peaks, ordinal
  patterns;
* bh_window_embedding, bh_window_lempel_ziv: correlation dimension, the
  Lyapunov estimates, Lempel-Ziv phrase counts.
Each calculator follows the reference's torch calculator term by term (the
1e-10 guards and the edge rules for short windows included; the calculator
is cited beside it).  On the CPU the same outputs come from torch (the CPU
template path).  A window holding a non-finite value gives NaN (inf for
lyapunov_e, as the reference) for every feature here (the extractor then
imputes them), where the reference's counting calculators would count its
finite entries.

Differences of rounding only: the reference reduces in the trajectory dtype
with torch's summation order, these sums run in double; counts that compare
a sample with the mean, a quantile or r standard deviations can differ for a
sample within rounding of the threshold.
"""

from __future__ import annotations

import math

import torch

from . import _lib

WF = _lib.WF

# the calculators built here with the parameter grids of the reference's
# comprehensive set (settings.py:223-340); None = no parameters.  Other
# parameter values of the count / crossing / lag / quantile families are not
# reduced by the kernel and are rejected.
COMPREHENSIVE_FEATURES: dict[str, list[dict] | None] = {
    "sum_values": None,
    "length": None,
    "abs_energy": None,
    "kurtosis": None,
    "skewness": None,
    "variation_coefficient": None,
    "absolute_sum_of_changes": None,
    "mean_abs_change": None,
    "mean_change": None,
    "mean_second_derivative_central": None,
    "count_above_mean": None,
    "count_below_mean": None,
    "count_above": [{"t": 0}],
    "count_below": [{"t": 0}],
    "has_duplicate_max": None,
    "has_duplicate_min": None,
    "has_variance_larger_than_standard_deviation": None,
    "large_standard_deviation": [{"r": r * 0.05} for r in range(1, 20)],
    "first_location_of_maximum": None,
    "first_location_of_minimum": None,
    "last_location_of_maximum": None,
    "last_location_of_minimum": None,
    "index_mass_quantile": [{"q": q} for q in [0.1, 0.2, 0.3, 0.4, 0.6, 0.7, 0.8, 0.9]],
    "longest_strike_above_mean": None,
    "longest_strike_below_mean": None,
    "number_crossing_m": [{"m": 0}, {"m": -1}, {"m": 1}],
    "number_peaks": [{"n": n} for n in [1, 3, 5, 10, 50]],
    "autocorrelation": [{"lag": lag} for lag in range(10)],
    "cid_ce": [{"normalize": True}, {"normalize": False}],
    "ratio_beyond_r_sigma": [{"r": x} for x in [0.5, 1, 1.5, 2, 2.5, 3, 5, 6, 7, 10]],
    "c3": [{"lag": lag} for lag in range(1, 4)],
    "time_reversal_asymmetry_statistic": [{"lag": lag} for lag in range(1, 4)],
    "count_value": [{"value": v} for v in [0, 1, -1]],
    "count_in_range": [{"min_val": -1, "max_val": 1}, {"min_val": -1e12, "max_val": 0},
                       {"min_val": 0, "max_val": 1e12}],
    "energy_ratio_by_chunks": [{"num_segments": 10, "segment_focus": i} for i in range(10)],
    "mean_n_absolute_max": [{"number_of_maxima": n} for n in [3, 5, 7]],
    "linear_trend": [{"attr": a} for a in ["pvalue", "rvalue", "intercept", "slope", "stderr"]],
    "linear_trend_timewise": [{"attr": a} for a in ["pvalue", "rvalue", "intercept", "slope",
                                                    "stderr"]],
    "partial_autocorrelation": [{"lag": lag} for lag in range(10)],
    "binned_entropy": [{"max_bins": 10}],
    "symmetry_looking": [{"r": r * 0.05} for r in range(20)],
}
# the calculators of the same set formed from the rfft of the window
# (bh_window_spectrum): complex bins 0 .. 99 and the aggregated power spectrum
SPECTRUM_FEATURES: dict[str, list[dict]] = {
    "fft_coefficient": [{"coeff": k, "attr": a} for a in ["real", "imag", "abs", "angle"]
                        for k in range(100)],
    "fft_aggregated": [{"aggtype": s} for s in ["centroid", "variance", "skew", "kurtosis"]],
    "spkt_welch_density": [{"coeff": c} for c in [2, 5, 8]],
    "fourier_entropy": [{"bins": x} for x in [2, 3, 5, 10, 100]],
}
COMPREHENSIVE_FEATURES.update(SPECTRUM_FEATURES)
# the calculators formed from the sorted window (bh_window_sorted)
_CQ_LEVELS = [0.0, 0.2, 0.4, 0.6, 0.8, 1.0]
_CQ_CORRIDORS = [(a, b) for a in range(5) for b in range(a + 1, 6)]
SORTED_FEATURES: dict[str, list[dict] | None] = {
    "has_duplicate": None,
    "ratio_value_number_to_time_series_length": None,
    "percentage_of_reoccurring_datapoints_to_all_datapoints": None,
    "percentage_of_reoccurring_values_to_all_values": None,
    "sum_of_reoccurring_data_points": None,
    "sum_of_reoccurring_values": None,
    "change_quantiles": [{"ql": ql, "qh": qh, "isabs": b, "f_agg": f}
                         for ql in [0.0, 0.2, 0.4, 0.6, 0.8] for qh in [0.2, 0.4, 0.6, 0.8, 1.0]
                         for b in [False, True] for f in ["mean", "var"] if ql < qh],
}
COMPREHENSIVE_FEATURES.update(SORTED_FEATURES)
# the calculators formed from the autocovariances of many lags (bh_window_autocov)
AUTOCOV_FEATURES: dict[str, list[dict]] = {
    "agg_autocorrelation": [{"f_agg": s, "maxlag": 40} for s in ["mean", "median", "var"]],
    "ar_coefficient": [{"coeff": c, "k": 10} for c in range(11)],
}
COMPREHENSIVE_FEATURES.update(AUTOCOV_FEATURES)
# the calculators of bh_window_trends: chunk trends, cubic fits, leading
# digits, wavelet coefficients
_ALT_SIZES, _ALT_AGGS = [5, 10, 50], ["max", "min", "mean", "var"]
_ALT_ATTRS = ["rvalue", "intercept", "slope", "stderr"]
_CWT_WIDTHS = [2, 5, 10, 20]
TRENDS_FEATURES: dict[str, list[dict] | None] = {
    "agg_linear_trend": [{"attr": a, "chunk_size": i, "f_agg": f} for a in _ALT_ATTRS
                         for i in _ALT_SIZES for f in _ALT_AGGS],
    "friedrich_coefficients": [{"coeff": c, "m": 3, "r": 30} for c in range(4)],
    "max_langevin_fixed_point": [{"m": 3, "r": 30}],
    "benford_correlation": None,
    "cwt_coefficients": [{"widths": (w,), "coeff": c, "w": w} for w in _CWT_WIDTHS
                         for c in range(15)],
}
COMPREHENSIVE_FEATURES.update(TRENDS_FEATURES)
# peaks of wide support and ordinal patterns (bh_window_peak_counts,
# bh_window_perm_entropy); number_peaks with n other than 1, 3, 5 joins them
PEAK_FEATURES: dict[str, list[dict]] = {
    "number_cwt_peaks": [{"max_width": n} for n in [1, 5]],
    "permutation_entropy": [{"tau": 1, "dimension": x} for x in [3, 4, 5, 6, 7]],
}
COMPREHENSIVE_FEATURES.update(PEAK_FEATURES)
_BLOCK_PEAKS = (1, 3, 5)
# delay-embedding estimates (bh_window_embedding)
EMBED_FEATURES: dict[str, list[dict]] = {
    "correlation_dimension": [{"emb_dim": 4, "lag": 1, "n_rvals": 50}],
    "lyapunov_r": [{"emb_dim": 10, "lag": 1, "trajectory_len": 20, "tau": 1.0}],
    "lyapunov_e": [{"emb_dim": 10, "matrix_dim": 4, "min_nb": 8, "min_tsep": 0, "tau": 1.0}],
}
# calculators giving several columns (the reference appends __0, __1, ..)
MULTI_OUTPUT = {"lyapunov_e": 4}
COMPREHENSIVE_FEATURES.update(EMBED_FEATURES)
# phrase counts (bh_window_lempel_ziv)
LZ_FEATURES: dict[str, list[dict]] = {
    "lempel_ziv_complexity": [{"bins": x} for x in [2, 3, 5, 10, 100]],
}
COMPREHENSIVE_FEATURES.update(LZ_FEATURES)
TR = dict(ALT=0, FRIED=48, LANGEVIN=52, BENFORD=53, CWT=54)
BH_TR_SLOTS = 114
# the remaining regression calculator formed from the window-feature slots
COMPREHENSIVE_FEATURES["augmented_dickey_fuller"] = [{"attr": a} for a in ["teststat", "pvalue",
                                                                          "usedlag"]]
MAX_AUTOCOV_LAG = 4095
# the largest rfft bin index bh_window_spectrum writes
MAX_SPECTRUM_COEFF = 4095
# calculators whose parameter may take any value (formed from the slots by
# a comparison with the parameter)
_FREE_PARAMS = {"large_standard_deviation", "symmetry_looking", "linear_trend",
                "linear_trend_timewise"}
_SIGMAS = [0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 5.0, 6.0, 7.0, 10.0]
_MASS_QS = [0.1, 0.2, 0.3, 0.4, 0.6, 0.7, 0.8, 0.9]


def check_params(name: str, params: dict | None) -> None:
    """ValueError unless the kernel's slots cover this parameter set."""
    grid = COMPREHENSIVE_FEATURES[name]
    if grid is None:
        if params:
            raise ValueError(f"{name} takes no parameters, got {params}")
        return
    if not params:
        raise ValueError(f"{name} needs parameters, one of {grid}")
    if name in _FREE_PARAMS:
        if name.startswith("linear_trend") and params.get("attr") not in (
                "pvalue", "rvalue", "intercept", "slope", "stderr"):
            raise ValueError(f"linear_trend: unknown attr {params.get('attr')!r}")
        return
    if name == "mean_n_absolute_max" and 1 <= int(params.get("number_of_maxima", 0)) <= 7:
        return
    if name == "number_peaks":
        if isinstance(params.get("n"), int) and 1 <= params["n"] <= 64:
            return
        raise ValueError(f"number_peaks: n must be an integer in [1, 64], got {params}")
    if name == "number_cwt_peaks":
        if isinstance(params.get("max_width"), int) and 1 <= params["max_width"] <= 8:
            return
        raise ValueError(f"number_cwt_peaks: max_width must be in [1, 8], got {params}")
    if name == "lempel_ziv_complexity":
        if isinstance(params.get("bins"), int) and 1 <= params["bins"] <= 255:
            return
        raise ValueError(f"lempel_ziv_complexity: bins must be an integer in [1, 255], got {params}")
    if name == "correlation_dimension":
        if (params.get("emb_dim", 4), params.get("lag", 1), params.get("n_rvals", 50)) == (4, 1, 50):
            return
        raise ValueError(f"correlation_dimension: emb_dim 4, lag 1, n_rvals 50 are built, got {params}")
    if name == "lyapunov_e":
        if (params.get("emb_dim", 10), params.get("matrix_dim", 4), params.get("min_nb", 8),
                params.get("min_tsep", 0)) == (10, 4, 8, 0) and float(params.get("tau", 1.0)) > 0:
            return
        raise ValueError("lyapunov_e: emb_dim 10, matrix_dim 4, min_nb 8, min_tsep 0 are built, "
                         f"got {params}")
    if name == "lyapunov_r":
        if (params.get("emb_dim", 10), params.get("lag", 1), params.get("trajectory_len", 20)) == \
                (10, 1, 20) and float(params.get("tau", 1.0)) > 0:
            return
        raise ValueError(f"lyapunov_r: emb_dim 10, lag 1, trajectory_len 20 are built, got {params}")
    if name == "permutation_entropy":
        if params.get("tau", 1) == 1 and params.get("dimension") in (3, 4, 5, 6, 7):
            return
        raise ValueError(f"permutation_entropy: tau 1 and dimension 3..7, got {params}")
    if name in ("autocorrelation", "partial_autocorrelation") and 0 <= int(params.get("lag", -1)) <= 9:
        return
    if name == "energy_ratio_by_chunks" and int(params.get("num_segments", 0)) == 10 and \
            0 <= int(params.get("segment_focus", -1)) < 10:
        return
    if name == "cid_ce" and "normalize" in params:
        return
    if name in ("fft_coefficient", "spkt_welch_density"):
        k = params.get("coeff", 0)
        if not (isinstance(k, int) and 0 <= k <= MAX_SPECTRUM_COEFF):
            raise ValueError(f"{name}: coeff must be an integer in [0, {MAX_SPECTRUM_COEFF}], "
                             f"got {k!r}")
        return
    if name in ("fft_aggregated", "fourier_entropy"):
        return
    if name == "agg_autocorrelation":
        if isinstance(params.get("maxlag", 40), int) and 1 <= params.get("maxlag", 40) <= MAX_AUTOCOV_LAG:
            return
        raise ValueError(f"agg_autocorrelation: maxlag must be in [1, {MAX_AUTOCOV_LAG}]")
    if name == "ar_coefficient":
        k, c = params.get("k", 1), params.get("coeff", 0)
        if isinstance(k, int) and isinstance(c, int) and 1 <= k <= 64 and c >= 0:
            return
        raise ValueError("ar_coefficient: k must be in [1, 64] and coeff >= 0")
    if name == "augmented_dickey_fuller":
        return
    if name == "agg_linear_trend":
        if params.get("chunk_size") in _ALT_SIZES and params.get("f_agg") in _ALT_AGGS and \
                params.get("attr") in _ALT_ATTRS:
            return
        raise ValueError(f"agg_linear_trend: chunk_size in {_ALT_SIZES}, f_agg in {_ALT_AGGS}, "
                         f"attr in {_ALT_ATTRS}; got {params}")
    if name in ("friedrich_coefficients", "max_langevin_fixed_point"):
        if int(params.get("m", 3)) == 3 and int(params.get("coeff", 0)) >= 0:
            return
        raise ValueError(f"{name}: the cubic (m = 3) is fitted, got {params}")
    if name == "cwt_coefficients":
        if params.get("w") in _CWT_WIDTHS and 0 <= int(params.get("coeff", -1)) < 15:
            return
        raise ValueError(f"cwt_coefficients: w in {_CWT_WIDTHS} and coeff in 0..14, got {params}")
    if name == "change_quantiles":
        if params.get("ql") in _CQ_LEVELS and params.get("qh") in _CQ_LEVELS and \
                params["ql"] < params["qh"] and params.get("f_agg", "mean") in ("mean", "var"):
            return
        raise ValueError(f"change_quantiles: ql < qh must be levels {_CQ_LEVELS} and f_agg "
                         f"mean or var, got {params}")
    if params not in grid:
        raise ValueError(f"{name}: parameters {params} are not reduced by the window kernel; "
                         f"supported: {grid}")


def _block_cpu(w: torch.Tensor) -> torch.Tensor:
    """The slots of bh_window_features for a CPU window (N, B, S), in torch."""
    x = w.to(torch.float64)
    N, B, S = x.shape
    out = torch.full((_lib.BH_WF_SLOTS, B, S), float("nan"), dtype=torch.float64)
    ok = torch.isfinite(x).all(0)
    ar = torch.arange(N, dtype=torch.float64).view(N, 1, 1)
    mx, mn = x.max(0).values, x.min(0).values
    o = {}
    o["N"] = torch.full((B, S), float(N), dtype=torch.float64)
    o["SUM"] = x.sum(0)
    o["SUMSQ"] = (x * x).sum(0)
    o["MIN"], o["MAX"] = mn, mx
    big = torch.full_like(x, float(N))
    o["FIRST_MAX"] = torch.where(x == mx, ar.expand_as(x), big).min(0).values
    o["LAST_MAX"] = torch.where(x == mx, ar.expand_as(x), -big).max(0).values
    o["FIRST_MIN"] = torch.where(x == mn, ar.expand_as(x), big).min(0).values
    o["LAST_MIN"] = torch.where(x == mn, ar.expand_as(x), -big).max(0).values
    dx = x[1:] - x[:-1]
    o["ABS_CHANGE"] = dx.abs().sum(0)
    o["SQ_CHANGE"] = (dx * dx).sum(0)
    o["X_FIRST"], o["X_LAST"] = x[0], x[-1]
    nanv = torch.full((B, S), float("nan"), dtype=torch.float64)
    o["X_SECOND"] = x[1] if N > 1 else nanv
    o["X_SECOND_LAST"] = x[-2] if N > 1 else nanv
    o["COUNT_POS"] = (x > 0).double().sum(0)
    o["COUNT_NEG"] = (x < 0).double().sum(0)
    o["COUNT_ZERO"] = (x == 0).double().sum(0)
    o["COUNT_ONE"] = (x == 1).double().sum(0)
    o["COUNT_MINUS_ONE"] = (x == -1).double().sum(0)
    o["RANGE_UNIT"] = ((x >= -1) & (x <= 1)).double().sum(0)
    o["RANGE_NONPOS"] = ((x >= -1e12) & (x <= 0)).double().sum(0)
    o["RANGE_NONNEG"] = ((x >= 0) & (x <= 1e12)).double().sum(0)
    for key, m in (("CROSS_0", 0.0), ("CROSS_M1", -1.0), ("CROSS_P1", 1.0)):
        a = x > m
        o[key] = (a[1:] != a[:-1]).double().sum(0)
    slots = {k: v for k, v in o.items()}
    seg = N // 10
    for j in range(10):
        slots[("CHUNK", j)] = (x[j * seg:(j + 1) * seg] ** 2).sum(0) if seg > 0 else torch.zeros(B, S, dtype=torch.float64)
    for lag in (1, 2, 3):
        if 2 * lag < N:
            a, b, c = x[:-2 * lag], x[lag:-lag], x[2 * lag:]
            slots[("C3", lag - 1)] = (a * b * c).sum(0)
            slots[("TREV", lag - 1)] = (c * c * b - b * a * a).sum(0)
        else:
            slots[("C3", lag - 1)] = slots[("TREV", lag - 1)] = torch.zeros(B, S, dtype=torch.float64)
    ax = x.abs()
    slots["SUM_ABS"] = ax.sum(0)
    k = min(7, N)
    top = torch.full((7, B, S), -1.0, dtype=torch.float64)
    top[:k] = ax.topk(k, dim=0).values
    for j in range(7):
        slots[("TOP_ABS", j)] = top[j]
    for key, n in (("PEAKS_1", 1), ("PEAKS_3", 3), ("PEAKS_5", 5)):
        if 2 * n >= N:
            slots[key] = torch.zeros(B, S, dtype=torch.float64)
            continue
        u = x.unfold(0, 2 * n + 1, 1)
        c = u[..., n]
        pk = (c > u[..., :n].max(-1).values) & (c > u[..., n + 1:].max(-1).values)
        slots[key] = pk.double().sum(0)
    mu = slots["SUM"] / N
    sh = x - x[0:1]
    m1 = sh.sum(0) / N
    sd = torch.sqrt(torch.clamp((sh * sh).sum(0) / N - m1 * m1, min=0.0))
    d = x - mu
    slots["CM2"] = (d * d).sum(0)
    slots["CM3"] = (d ** 3).sum(0)
    slots["CM4"] = (d ** 4).sum(0)
    slots["SS_TX"] = ((ar - 0.5 * (N - 1)) * d).sum(0)
    slots["COUNT_ABOVE_MEAN"] = (x > mu).double().sum(0)
    slots["COUNT_BELOW_MEAN"] = (x < mu).double().sum(0)
    for key, mask in (("STRIKE_ABOVE", x > mu), ("STRIKE_BELOW", x < mu)):
        run = torch.zeros(B, S, dtype=torch.float64)
        best = torch.zeros(B, S, dtype=torch.float64)
        for i in range(N):
            run = (run + 1) * mask[i].double()
            best = torch.maximum(best, run)
        slots[key] = best
    slots["COUNT_MAX"] = (x == mx).double().sum(0)
    slots["COUNT_MIN"] = (x == mn).double().sum(0)
    for j, r in enumerate(_SIGMAS):
        slots[("BEYOND", j)] = (d.abs() > r * sd).double().sum(0)
    for lag in range(1, 10):
        slots[("AUTOCOV", lag - 1)] = (d[:-lag] * d[lag:]).sum(0) if lag < N else torch.zeros(B, S, dtype=torch.float64)
    cum = ax.cumsum(0)
    for j, q in enumerate(_MASS_QS):
        hit = cum >= q * slots["SUM_ABS"]
        slots[("MASS", j)] = torch.where(hit.any(0), hit.double().argmax(0).double(),
                                         torch.full((B, S), float(N - 1), dtype=torch.float64))
    # binned_entropy bins in the series dtype (torch_features_entropy_complexity.py:58-94)
    wmn, wmx = w.min(0).values, w.max(0).values
    rng = wmx - wmn
    rng = torch.where(rng == 0, torch.ones_like(rng), rng)
    bins = ((w - wmn) / (rng + 1e-10) * (10 - 1e-6)).long().clamp(0, 9)
    for j in range(10):
        slots[("BINS", j)] = (bins == j).double().sum(0)
    for key, v in slots.items():
        idx = WF[key] if isinstance(key, str) else WF[key[0]] + key[1]
        out[idx] = v
    return torch.where(ok.unsqueeze(0), out, torch.full_like(out, float("nan")))


def window_feature_block(y: torch.Tensor, start: int) -> torch.Tensor:
    """(BH_WF_SLOTS, B, S) float64 slots of the window rows [start, N) of a
    kept trajectory tensor y (N, B, S): bh_window_features on the device, the
    torch restatement on the CPU."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        return _block_cpu(y[start:])
    y = y.contiguous()
    out = torch.empty((_lib.BH_WF_SLOTS, B, S), dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_features(dt, y.data_ptr(), N, B, S, int(start),
                                                 out.data_ptr(), s.cuda_stream),
                   "bh_window_features")
    return out


def _pacf_all(blk: torch.Tensor, n: float) -> list:
    """Partial autocorrelations at lags 1..min(9, n - 1) by one Durbin-Levinson
    recursion (torch_features_autocorrelation.py:31-69): entry k is the
    coefficient phi_k of step k."""
    g = lambda key, j=0: blk[WF[key] + j]  # noqa: E731
    maxlag = min(9, int(n) - 1)
    out = [None]
    if maxlag < 1:
        return out
    acf = [None] + [g("AUTOCOV", l - 1) / (g("CM2") + 1e-10) for l in range(1, maxlag + 1)]
    phi = [None, acf[1]]
    out.append(acf[1])
    for k in range(2, maxlag + 1):
        num = acf[k] - sum(phi[j] * acf[k - j] for j in range(1, k))
        den = 1.0 - sum(phi[j] * acf[j] for j in range(1, k))
        pk = num / (den + 1e-10)
        phi = [None] + [phi[j] - pk * phi[k - j] for j in range(1, k)] + [pk]
        out.append(pk)
    return out


def comprehensive_feature(name: str, params: dict | None, blk: torch.Tensor, n_win: int,
                          median: torch.Tensor | None = None,
                          cache: dict | None = None) -> torch.Tensor:
    """Feature `name` (B, S) float64 from the slots of windows of n_win rows
    (the reference calculator it restates is cited beside each).
    symmetry_looking needs the window median.  NaN where the window holds a
    non-finite value (its N slot is NaN).  cache: a dict kept across the
    calls on one block (the partial autocorrelations of all lags come from
    one recursion)."""
    p = params or {}
    g = lambda key, j=0: blk[WF[key] + j]  # noqa: E731
    N = g("N")
    n = float(n_win)
    mean = g("SUM") / N
    m2 = g("CM2") / N
    sd = torch.sqrt(m2)
    rng = g("MAX") - g("MIN")
    if name == "sum_values":  # statistical.py:20-26
        return g("SUM")
    if name == "length":  # statistical.py:50-57
        return N
    if name == "abs_energy":  # statistical.py:119-121
        return g("SUMSQ")
    if name == "kurtosis":  # statistical.py:125-135
        g2 = (g("CM4") / N) / (m2 ** 2 + 1e-10) - 3
        return ((n - 1) / ((n - 2) * (n - 3) + 1e-10)) * ((n + 1) * g2 + 6)
    if name == "skewness":  # statistical.py:139-150
        g1 = (g("CM3") / N) / (m2 ** 1.5 + 1e-10)
        return g1 * math.sqrt(max(n * (n - 1), 0.0)) / (n - 2 + 1e-10)
    if name == "variation_coefficient":  # statistical.py:160-162
        return sd / (mean.abs() + 1e-10)
    if name == "absolute_sum_of_changes":  # change.py
        return g("ABS_CHANGE")
    if name == "mean_abs_change":
        return g("ABS_CHANGE") / (N - 1)
    if name == "mean_change":
        return (g("X_LAST") - g("X_FIRST")) / (N - 1)
    if name == "mean_second_derivative_central":
        return (g("X_LAST") - g("X_SECOND_LAST") - g("X_SECOND") + g("X_FIRST")) / (2 * (N - 2))
    if name == "count_above_mean":  # count.py
        return g("COUNT_ABOVE_MEAN")
    if name == "count_below_mean":
        return g("COUNT_BELOW_MEAN")
    if name == "count_above":
        return g("COUNT_POS") / N
    if name == "count_below":
        return g("COUNT_NEG") / N
    if name == "count_value":
        return g({0: "COUNT_ZERO", 1: "COUNT_ONE", -1: "COUNT_MINUS_ONE"}[int(p["value"])])
    if name == "count_in_range":
        key = {(-1, 1): "RANGE_UNIT", (-1e12, 0): "RANGE_NONPOS", (0, 1e12): "RANGE_NONNEG"}
        return g(key[(p["min_val"], p["max_val"])])
    if name == "has_duplicate_max":  # boolean.py
        return (g("COUNT_MAX") > 1).double().where(~N.isnan(), N)
    if name == "has_duplicate_min":
        return (g("COUNT_MIN") > 1).double().where(~N.isnan(), N)
    if name == "has_variance_larger_than_standard_deviation":
        return (sd > 1).double().where(~N.isnan(), N)
    if name == "large_standard_deviation":
        return (sd > float(p["r"]) * rng).double().where(~N.isnan(), N)
    if name == "first_location_of_maximum":  # location.py
        return g("FIRST_MAX") / N
    if name == "first_location_of_minimum":
        return g("FIRST_MIN") / N
    if name == "last_location_of_maximum":
        return g("LAST_MAX") / N
    if name == "last_location_of_minimum":
        return g("LAST_MIN") / N
    if name == "index_mass_quantile":
        return g("MASS", _MASS_QS.index(float(p["q"]))) / N
    if name == "longest_strike_above_mean":  # pattern.py
        return g("STRIKE_ABOVE")
    if name == "longest_strike_below_mean":
        return g("STRIKE_BELOW")
    if name == "number_crossing_m":
        return g({0: "CROSS_0", -1: "CROSS_M1", 1: "CROSS_P1"}[int(p["m"])])
    if name == "number_peaks":
        return g({1: "PEAKS_1", 3: "PEAKS_3", 5: "PEAKS_5"}[int(p["n"])])
    if name == "autocorrelation":  # autocorrelation.py:13-28
        lag = int(p["lag"])
        if lag == 0:
            return torch.ones_like(N).where(~N.isnan(), N)
        if lag >= n:
            return torch.zeros_like(N).where(~N.isnan(), N)
        return g("AUTOCOV", lag - 1) / (g("CM2") + 1e-10)
    if name == "cid_ce":  # entropy_complexity.py:182-190
        ce = torch.sqrt(g("SQ_CHANGE"))
        return ce / (sd + 1e-10) if p.get("normalize", True) else ce
    if name == "ratio_beyond_r_sigma":  # statistical.py:175-181
        return g("BEYOND", _SIGMAS.index(float(p["r"]))) / N
    if name in ("c3", "time_reversal_asymmetry_statistic"):  # advanced.py:84-123
        lag = int(p["lag"])
        if 2 * lag >= n:
            return torch.zeros_like(N).where(~N.isnan(), N)
        return g("C3" if name == "c3" else "TREV", lag - 1) / (N - 2 * lag)
    if name == "energy_ratio_by_chunks":  # advanced.py:93-106
        if int(n) // 10 == 0:
            return torch.zeros_like(N).where(~N.isnan(), N)
        return g("CHUNK", int(p["segment_focus"])) / (g("SUMSQ") + 1e-10)
    if name == "mean_n_absolute_max":  # statistical.py:165-172
        k = min(int(p["number_of_maxima"]), int(n))
        return sum(g("TOP_ABS", j) for j in range(k)) / k
    if name == "partial_autocorrelation":  # autocorrelation.py:31-69 (Durbin-Levinson)
        lag = int(p["lag"])
        if lag == 0:
            return torch.ones_like(N).where(~N.isnan(), N)
        if lag > int(n) - 1:
            return torch.zeros_like(N).where(~N.isnan(), N)
        # the coefficient phi_lag of step lag is the same whether the
        # recursion stops there (the reference) or runs on to lag 9
        pacf = cache.get("pacf") if cache is not None else None
        if pacf is None:
            pacf = _pacf_all(blk, n)
            if cache is not None:
                cache["pacf"] = pacf
        return pacf[lag]
    if name == "binned_entropy":  # entropy_complexity.py:58-94
        ent = torch.zeros_like(N)
        for j in range(10):
            c = g("BINS", j)
            pr = c / N
            ent = ent - torch.where(c > 0, pr * torch.log(pr.clamp(min=1e-10)), torch.zeros_like(pr))
        return torch.where(g("MAX") == g("MIN"), torch.zeros_like(ent), ent).where(~N.isnan(), N)
    if name in ("linear_trend", "linear_trend_timewise"):  # trend.py:11-58
        tbar = 0.5 * (n - 1)
        ss_tt = n * (n * n - 1) / 12.0
        ss_tx = g("SS_TX")
        ss_xx = g("CM2")
        slope = ss_tx / (ss_tt + 1e-10)
        attr = p.get("attr", "slope")
        if attr == "slope":
            return slope
        if attr == "intercept":
            return mean - slope * tbar
        rvalue = ss_tx / (torch.sqrt(ss_tt * ss_xx) + 1e-10)
        if attr == "rvalue":
            return rvalue
        if attr == "pvalue":
            # sqrt(n - 2) of a window of one row is NaN, as in the reference
            t_stat = rvalue * (math.sqrt(n - 2) if n >= 2 else math.nan) / (
                torch.sqrt(1 - rvalue ** 2) + 1e-10)
            return 1 / (1 + t_stat.abs())
        resid = torch.clamp(ss_xx - 2 * slope * ss_tx + slope * slope * ss_tt, min=0.0)
        return torch.sqrt(resid / (n - 2) / (ss_tt + 1e-10))
    if name == "augmented_dickey_fuller":  # trend.py:134-172: regression of the changes
        # on the lagged values, from the slots: with dl = x_last - mean, the
        # lagged values' centred square sum is CM2 - dl^2 - dl^2 / (n - 1) and
        # their cross sum with the changes AUTOCOV_1 - (CM2 - dl^2) + dl / (n - 1)
        # (x_last - x_first); the changes' own is SQ_CHANGE - (x_last - x_first)^2 / (n - 1)
        if p.get("attr", "teststat") == "usedlag":
            return torch.ones_like(N).where(~N.isnan(), N)
        m = n - 1
        dl = g("X_LAST") - mean
        span = g("X_LAST") - g("X_FIRST")
        sl2 = g("CM2") - dl * dl
        ss_xx = sl2 - dl * dl / m
        ss_xy = g("AUTOCOV") - sl2 + dl / m * span
        ss_yy = g("SQ_CHANGE") - span * span / m
        beta = ss_xy / (ss_xx + 1e-10)
        resid = torch.clamp(ss_yy - 2 * beta * ss_xy + beta * beta * ss_xx, min=0.0)
        se = torch.sqrt(resid / (n - 3)) / (torch.sqrt(ss_xx) + 1e-10)
        t_stat = beta / (se + 1e-10)
        if p.get("attr", "teststat") == "pvalue":
            return 1 / (1 + t_stat.abs())
        return t_stat
    if name == "symmetry_looking":  # statistical.py:184-190
        if median is None:
            raise ValueError("symmetry_looking needs the window median")
        return ((mean - median.double()).abs() < float(p["r"]) * rng).double().where(~N.isnan(), N)
    raise KeyError(f"feature {name!r} is not a comprehensive window feature")


def spectrum_needs(plan) -> tuple[int, bool]:
    """(number of complex bins, whether the aggregated spectrum is needed)
    for the spectrum features of a plan of (state, name, params)."""
    n_coeff, agg = 0, False
    for _, name, params in plan:
        if name in ("fft_coefficient", "spkt_welch_density"):
            n_coeff = max(n_coeff, int((params or {}).get("coeff", 0)) + 1)
        elif name in ("fft_aggregated", "fourier_entropy"):
            agg = True
    return n_coeff, agg


def _spectrum_cpu(w: torch.Tensor, n_coeff: int, agg: bool):
    """The outputs of bh_window_spectrum for a CPU window (N, B, S), in torch."""
    X = torch.fft.rfft(w.to(torch.float64), dim=0)
    nb = X.shape[0]
    coeff = None
    if n_coeff:
        coeff = torch.zeros((n_coeff, 2) + tuple(w.shape[1:]), dtype=torch.float64)
        k = min(n_coeff, nb)
        coeff[:k, 0] = X[:k].real
        coeff[:k, 1] = X[:k].imag
        coeff[0, 1] = 0.0  # the DC and Nyquist bins are real (+0)
        if w.shape[0] % 2 == 0 and w.shape[0] // 2 < n_coeff:
            coeff[w.shape[0] // 2, 1] = 0.0
    aggs = None
    if agg:
        P = X.real ** 2 + X.imag ** 2
        pn = P / (P.sum(0) + 1e-10)
        b = torch.arange(nb, dtype=torch.float64).view(nb, 1, 1)
        cen = (b * pn).sum(0)
        d = b - cen
        var = (d ** 2 * pn).sum(0)
        pc = pn.clamp(min=1e-10)
        aggs = torch.stack([cen, var, (d ** 3 * pn).sum(0) / (var ** 1.5 + 1e-10),
                            (d ** 4 * pn).sum(0) / (var ** 2 + 1e-10),
                            -(pc * pc.log()).sum(0)])
    return coeff, aggs


def window_spectrum(y: torch.Tensor, start: int, n_coeff: int, agg: bool):
    """(coeff (n_coeff, 2, B, S) or None, agg (5, B, S) or None), float64:
    the rfft bins 0 .. n_coeff - 1 of each series of the window rows
    [start, N) of y (N, B, S) as (re, im) (0 past the last bin), and the
    centroid, variance, skew, kurtosis and entropy of its normalised power
    spectrum; bh_window_spectrum on the device, torch.fft on the CPU."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if not 0 <= n_coeff <= MAX_SPECTRUM_COEFF + 1:
        raise ValueError(f"n_coeff must be in [0, {MAX_SPECTRUM_COEFF + 1}]")
    if not n_coeff and not agg:
        return None, None
    if y.device.type != "cuda":
        return _spectrum_cpu(y[start:], n_coeff, agg)
    y = y.contiguous()
    coeff = (torch.empty((n_coeff, 2, B, S), dtype=torch.float64, device=y.device)
             if n_coeff else None)
    aggs = torch.empty((5, B, S), dtype=torch.float64, device=y.device) if agg else None
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_spectrum(
            dt, y.data_ptr(), N, B, S, int(start), int(n_coeff),
            coeff.data_ptr() if coeff is not None else None,
            aggs.data_ptr() if aggs is not None else None, s.cuda_stream),
            "bh_window_spectrum")
    return coeff, aggs


_AGGTYPES = {"centroid": 0, "variance": 1, "skew": 2, "kurtosis": 3}


def spectrum_feature(name: str, params: dict | None, coeff: torch.Tensor | None,
                     agg: torch.Tensor | None, n_win: int) -> torch.Tensor:
    """Spectrum feature `name` (B, S) float64 from the outputs of
    window_spectrum (torch_features_frequency.py:11-70,
    torch_features_entropy_complexity.py:97-107)."""
    p = params or {}
    if name == "fft_coefficient":  # frequency.py:11-28; an unknown attr gives abs
        k = int(p.get("coeff", 0))
        re, im = coeff[k, 0], coeff[k, 1]
        attr = p.get("attr", "abs")
        if attr == "real":
            return re
        if attr == "imag":
            return im
        if attr == "angle":
            return torch.atan2(im, re)
        return torch.hypot(re, im)
    if name == "spkt_welch_density":  # frequency.py:59-67
        k = int(p.get("coeff", 0))
        return (coeff[k, 0] ** 2 + coeff[k, 1] ** 2) / float(n_win)
    if name == "fft_aggregated":  # frequency.py:31-56; an unknown aggtype gives the centroid
        return agg[_AGGTYPES.get(p.get("aggtype", "centroid"), 0)]
    if name == "fourier_entropy":  # entropy_complexity.py:97-107; bins is not used
        return agg[4]
    raise KeyError(f"feature {name!r} is not a spectrum feature")


def _torch_quantile_sorted(srt: torch.Tensor, q: float) -> torch.Tensor:
    """torch.quantile's linear interpolation of sorted rows (N, ...) in
    their dtype (rank = q (N - 1) in that dtype, then lerp)."""
    N = srt.shape[0]
    rank = torch.tensor(q, dtype=srt.dtype) * (N - 1)
    lo = int(rank)
    hi = min(int(torch.ceil(rank)), N - 1)
    return torch.lerp(srt[lo], srt[hi], rank - lo)


def _sorted_cpu(w: torch.Tensor, cq: bool):
    """The outputs of bh_window_sorted for a CPU window (N, B, S), in torch."""
    N = w.shape[0]
    srt = w.sort(0).values
    x = srt.double()
    ok = torch.isfinite(x).all(0)
    eq = srt[1:] == srt[:-1]
    z = torch.zeros((1,) + tuple(w.shape[1:]), dtype=torch.bool)
    prev = torch.cat([z, eq])
    nxt = torch.cat([eq, z])
    grp = prev | nxt
    first = nxt & ~prev
    dup = torch.stack([eq.double().sum(0), grp.double().sum(0), first.double().sum(0),
                       (x * grp).sum(0), (x * first).sum(0)])
    dup = torch.where(ok, dup, torch.full_like(dup, float("nan")))
    med = torch.where(ok, srt[(N - 1) // 2], torch.full_like(srt[0], float("nan")))
    out_cq = None
    if cq:
        lv = [_torch_quantile_sorted(srt, q) for q in _CQ_LEVELS]
        d = (w[1:].double() - w[:-1].double())
        lo, hi = torch.minimum(w[1:], w[:-1]), torch.maximum(w[1:], w[:-1])
        rows = []
        for a, b in _CQ_CORRIDORS:
            m = ((lo >= lv[a]) & (hi <= lv[b])).double()
            c = m.sum(0)
            for v in (d, d.abs()):
                mean = (v * m).sum(0) / (c + 1e-10)
                var = ((v - mean) ** 2 * m).sum(0) / (c + 1e-10)
                rows += [torch.where(c > 0, mean, torch.zeros_like(mean)),
                         torch.where(c > 0, var, torch.zeros_like(var))]
        out_cq = torch.stack(rows)
        out_cq = torch.where(ok, out_cq, torch.full_like(out_cq, float("nan")))
    return dup, out_cq, med


def window_sorted(y: torch.Tensor, start: int, cq: bool):
    """(dup (5, B, S), cq (60, B, S) or None, median (B, S)): the
    sorted-window statistics of bh_window_sorted for the window rows
    [start, N) of y (N, B, S) (the kernel on the device, torch on the CPU).
    dup (float64) = adjacent equal pairs, values in groups of equal values,
    groups, sum of the values in groups, sum of one value per group; cq
    (float64) row 4 c + k = corridor c of _CQ_CORRIDORS, k = mean, var of the
    changes, mean, var of |changes|; median in the series dtype (the lower
    middle value, NaN for a non-finite window)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        return _sorted_cpu(y[start:], cq)
    y = y.contiguous()
    dup = torch.empty((5, B, S), dtype=torch.float64, device=y.device)
    levels = torch.empty((6, B, S), dtype=y.dtype, device=y.device) if cq else None
    out_cq = torch.empty((60, B, S), dtype=torch.float64, device=y.device) if cq else None
    med = torch.empty((B, S), dtype=y.dtype, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_sorted(
            dt, y.data_ptr(), N, B, S, int(start), dup.data_ptr(),
            levels.data_ptr() if levels is not None else None, med.data_ptr(), s.cuda_stream),
            "bh_window_sorted")
        if cq:
            _lib.check(_lib.lib().bh_window_change_quantiles(
                dt, y.data_ptr(), N, B, S, int(start), levels.data_ptr(), out_cq.data_ptr(),
                s.cuda_stream), "bh_window_change_quantiles")
    return dup, out_cq, med


def sorted_feature(name: str, params: dict | None, dup: torch.Tensor, cq: torch.Tensor | None,
                   n_win: int) -> torch.Tensor:
    """Sorted-window feature `name` (B, S) float64 from window_sorted's
    outputs (torch_features_boolean.py:11-19, torch_features_reocurrance.py:
    11-161, torch_features_change.py:20-139)."""
    n = float(n_win)
    D, R, G, SP, SV = dup
    if name == "has_duplicate":
        return (D > 0).double().where(~D.isnan(), D)
    if name == "ratio_value_number_to_time_series_length":
        return (n - D) / n
    if name == "percentage_of_reoccurring_datapoints_to_all_datapoints":
        return G / ((n - D) + 1e-10)
    if name == "percentage_of_reoccurring_values_to_all_values":
        return R / n
    if name == "sum_of_reoccurring_data_points":
        return SP
    if name == "sum_of_reoccurring_values":
        return SV
    if name == "change_quantiles":
        p = params or {}
        c = _CQ_CORRIDORS.index((_CQ_LEVELS.index(p["ql"]), _CQ_LEVELS.index(p["qh"])))
        k = (2 if p.get("isabs", True) else 0) + (1 if p.get("f_agg", "mean") == "var" else 0)
        return cq[4 * c + k]
    raise KeyError(f"feature {name!r} is not a sorted-window feature")


def autocov_needs(plan) -> int:
    """The largest lag the autocovariance features of a plan read."""
    lag = 0
    for _, name, params in plan:
        p = params or {}
        if name == "agg_autocorrelation":
            lag = max(lag, int(p.get("maxlag", 40)))
        elif name == "ar_coefficient":
            lag = max(lag, int(p.get("k", 1)))
    return lag


def _autocov_cpu(w: torch.Tensor, max_lag: int) -> torch.Tensor:
    """The outputs of bh_window_autocov for a CPU window (N, B, S), in torch."""
    x = w.double()
    N = x.shape[0]
    d = x - x.mean(0)
    out = torch.zeros((max_lag + 1,) + tuple(x.shape[1:]), dtype=torch.float64)
    for k in range(min(max_lag, N - 1) + 1):
        out[k] = (d[: N - k] * d[k:]).sum(0)
    return out


def window_autocov(y: torch.Tensor, start: int, max_lag: int) -> torch.Tensor:
    """(max_lag + 1, B, S) float64 autocovariances C_k of the window rows
    [start, N) of y (bh_window_autocov on the device, torch on the CPU)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if not 0 <= max_lag <= MAX_AUTOCOV_LAG:
        raise ValueError(f"max_lag must be in [0, {MAX_AUTOCOV_LAG}]")
    if y.device.type != "cuda":
        return _autocov_cpu(y[start:], max_lag)
    y = y.contiguous()
    out = torch.empty((max_lag + 1, B, S), dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_autocov(dt, y.data_ptr(), N, B, S, int(start), int(max_lag),
                                                out.data_ptr(), s.cuda_stream),
                   "bh_window_autocov")
    return out


def _yule_walker(t: list, k: int) -> list:
    """phi_1..phi_k of the symmetric Toeplitz system sum_j phi_j t_|i-j| = t_i
    (i = 1..k) by the Durbin-Levinson recursion, element-wise over series."""
    E = t[0]
    phi: list = []
    for m in range(1, k + 1):
        num = t[m] - sum((phi[j] * t[m - 1 - j] for j in range(m - 1)), torch.zeros_like(E))
        kap = num / E
        phi = [phi[j] - kap * phi[m - 2 - j] for j in range(m - 1)] + [kap]
        E = E * (1 - kap * kap)
    return phi


def autocov_feature(name: str, params: dict | None, C: torch.Tensor, n_win: int,
                    cache: dict | None = None) -> torch.Tensor:
    """Autocovariance feature `name` (B, S) float64 from window_autocov's
    output (torch_features_autocorrelation.py:72-96, torch_features_trend.py:
    86-131).  cache: a dict kept across calls on one window (the AR fit of
    one order serves all its coefficients)."""
    p = params or {}
    if name == "agg_autocorrelation":
        L = min(int(p.get("maxlag", 40)), n_win - 1)
        acf = C[1:L + 1] / (C[0] + 1e-10)
        f = p.get("f_agg", "mean")
        if f == "median":
            return acf.sort(0).values[(L - 1) // 2] if L >= 1 else torch.full_like(C[0], float("nan"))
        if f == "var":
            return acf.var(0, correction=0)
        return acf.mean(0)
    if name == "ar_coefficient":
        k, coeff = int(p.get("k", 1)), int(p.get("coeff", 0))
        if coeff >= k:  # coeff == k is zero in the reference as well as coeff > k
            return torch.where(torch.isfinite(C[0]), 0.0, float("nan")).to(C.dtype)
        key = ("ar", k)
        if cache is None or key not in cache:
            acf = C[: k + 1] / (C[0] + 1e-10)
            t = [acf[0] + 1e-6] + [acf[j] for j in range(1, k + 1)]
            phi = _yule_walker(t, k)
            if cache is None:
                return phi[coeff]
            cache[key] = phi
        return cache[key][coeff]
    raise KeyError(f"feature {name!r} is not an autocovariance feature")


def _trends_cpu(w: torch.Tensor, want_cwt: bool) -> torch.Tensor:
    """The outputs of bh_window_trends for a CPU window (N, B, S), in torch."""
    x = w.double()
    N, B, S = x.shape
    rows = BH_TR_SLOTS if want_cwt else TR["CWT"]
    out = torch.zeros((rows, B, S), dtype=torch.float64)
    ok = torch.isfinite(x).all(0)
    for si, cs in enumerate(_ALT_SIZES):
        nc = N // cs
        if nc < 2:
            continue
        ch = x[: nc * cs].reshape(nc, cs, B, S)
        aggs = [ch.max(1).values, ch.min(1).values, ch.mean(1), ch.var(1, correction=0)]
        t = torch.arange(nc, dtype=torch.float64).view(nc, 1, 1)
        tb = t.mean()
        ss_tt = ((t - tb) ** 2).sum()
        for fi, a in enumerate(aggs):
            am = a.mean(0)
            ss_tx = ((t - tb) * (a - am)).sum(0)
            ss_xx = ((a - am) ** 2).sum(0)
            slope = ss_tx / (ss_tt + 1e-10)
            res = ((a - (slope * t + (am - slope * tb))) ** 2).sum(0)
            base = TR["ALT"] + (si * 4 + fi) * 4
            out[base] = ss_tx / (torch.sqrt(ss_tt * ss_xx) + 1e-10)
            out[base + 1] = am - slope * tb
            out[base + 2] = slope
            out[base + 3] = torch.sqrt(res / (nc - 2) / (ss_tt + 1e-10))
    pos, vel = x[:-1], x[1:] - x[:-1]
    V = torch.stack([pos ** 3, pos ** 2, pos, torch.ones_like(pos)], -1)  # (N-1, B, S, 4)
    A = torch.einsum("nbsi,nbsj->bsij", V, V) + 1e-6 * torch.eye(4, dtype=torch.float64)
    rhs = torch.einsum("nbsi,nbs->bsi", V, vel)
    cf = torch.linalg.solve(torch.nan_to_num(A), torch.nan_to_num(rhs))  # (B, S, 4)
    for k in range(4):
        out[TR["FRIED"] + k] = cf[..., k]
    fit = (V * cf.unsqueeze(0)).sum(-1)  # (N-1, B, S)
    if N >= 3:
        chg = fit[:-1] * fit[1:] < 0
        best = torch.where(chg, pos[:-1], torch.full_like(pos[:-1], -math.inf)).max(0).values
        out[TR["LANGEVIN"]] = torch.where(best == -math.inf, torch.zeros_like(best), best)
    a = x.abs()
    nz = (a > 0) & torch.isfinite(a)
    aa = torch.where(nz, a, torch.ones_like(a))
    dg = torch.floor(aa / 10.0 ** torch.floor(torch.log10(aa))).long()  # 0 or 10: not counted
    cnt = torch.stack([((dg == d) & nz).double().sum(0) for d in range(1, 10)])  # (9, B, S)
    tot = cnt.sum(0)
    bf = torch.tensor([math.log10(1 + 1 / d) for d in range(1, 10)], dtype=torch.float64).view(9, 1, 1)
    fr = cnt / (tot + 1e-10)
    corr = ((fr - fr.mean(0)) * (bf - bf.mean())).sum(0) / (fr.std(0) * bf.std() * 9 + 1e-10)
    out[TR["BENFORD"]] = torch.where(tot >= 10, corr, torch.zeros_like(corr))
    if want_cwt:
        xs = x.permute(1, 2, 0).reshape(B * S, 1, N)
        for wi, wd in enumerate(_CWT_WIDTHS):
            t = torch.arange(-4 * wd, 4 * wd + 1, dtype=torch.float64)
            psi = (2 / (math.sqrt(3.0 * wd) * math.pi ** 0.25)) * (1 - (t / wd) ** 2) * \
                torch.exp(-t ** 2 / (2.0 * wd * wd))
            pad = 4 * wd
            padded = torch.nn.functional.pad(xs, (pad, pad), mode="reflect" if N > pad else "replicate")
            conv = torch.nn.functional.conv1d(padded, psi.view(1, 1, -1))[:, 0]  # (B*S, N)
            for c in range(15):
                out[TR["CWT"] + wi * 15 + c] = conv[:, c].reshape(B, S) if c < N else 0.0
    return torch.where(ok, out, torch.full_like(out, float("nan")))


def window_trends(y: torch.Tensor, start: int, want_cwt: bool) -> torch.Tensor:
    """(BH_TR_SLOTS or TR["CWT"] rows, B, S) float64 outputs of
    bh_window_trends for the window rows [start, N) of y (the kernel on the
    device, torch on the CPU)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        return _trends_cpu(y[start:], want_cwt)
    y = y.contiguous()
    rows = BH_TR_SLOTS if want_cwt else TR["CWT"]
    out = torch.empty((rows, B, S), dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_trends(dt, y.data_ptr(), N, B, S, int(start), int(want_cwt),
                                               out.data_ptr(), s.cuda_stream), "bh_window_trends")
    return out


def trends_feature(name: str, params: dict | None, out: torch.Tensor) -> torch.Tensor:
    """Feature `name` (B, S) float64 from window_trends' output
    (torch_features_trend.py:60-83, torch_features_dynamical.py:87-177,
    torch_features_advanced.py:21-78, torch_features_frequency.py:71-122)."""
    p = params or {}
    nan_or = lambda v: torch.where(torch.isnan(out[TR["BENFORD"]]), out[TR["BENFORD"]], v)  # noqa: E731
    if name == "agg_linear_trend":
        k = (_ALT_SIZES.index(p["chunk_size"]) * 4 + _ALT_AGGS.index(p["f_agg"])) * 4 + \
            _ALT_ATTRS.index(p["attr"])
        return out[TR["ALT"] + k]
    if name == "friedrich_coefficients":
        c = int(p.get("coeff", 0))
        return out[TR["FRIED"] + c] if c < 4 else nan_or(torch.zeros_like(out[0]))
    if name == "max_langevin_fixed_point":
        return out[TR["LANGEVIN"]]
    if name == "benford_correlation":
        return out[TR["BENFORD"]]
    if name == "cwt_coefficients":
        w = p.get("w", 2)
        if w not in tuple(p.get("widths", (2,))):
            return nan_or(torch.zeros_like(out[0]))
        return out[TR["CWT"] + _CWT_WIDTHS.index(w) * 15 + int(p.get("coeff", 0))]
    raise KeyError(f"feature {name!r} is not a chunk-trend / fit feature")


def _peak_counts_cpu(w: torch.Tensor, supports: list) -> torch.Tensor:
    """The outputs of bh_window_peak_counts for a CPU window (N, B, S)."""
    x = w.double()
    N, B, S = x.shape
    out = torch.zeros((12, B, S), dtype=torch.float64)
    for r, k in enumerate(supports):
        if k <= 0 or 2 * k >= N:
            continue
        u = x.unfold(0, 2 * k + 1, 1)  # (N - 2k, B, S, 2k + 1)
        c = u[..., k]
        out[r] = ((c > u[..., :k].max(-1).values) & (c > u[..., k + 1:].max(-1).values)).double().sum(0)
    for wd in range(1, 9):
        if 2 * wd >= N:
            continue
        u = x.unfold(0, 2 * wd + 1, 1)
        out[4 + wd - 1] = (u[..., wd] == u.max(-1).values).double().sum(0)
    ok = torch.isfinite(x).all(0)
    return torch.where(ok, out, torch.full_like(out, float("nan")))


def window_peak_counts(y: torch.Tensor, start: int, supports: list) -> torch.Tensor:
    """(12, B, S) float64: strict peak counts for up to four supports (rows
    0..3) and the per-width counts of number_cwt_peaks (rows 4..11) of the
    window rows [start, N) of y (bh_window_peak_counts on the device)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    sup = (list(supports) + [0, 0, 0, 0])[:4]
    if y.device.type != "cuda":
        return _peak_counts_cpu(y[start:], sup)
    y = y.contiguous()
    out = torch.empty((12, B, S), dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_peak_counts(dt, y.data_ptr(), N, B, S, int(start), *sup,
                                                    out.data_ptr(), s.cuda_stream),
                   "bh_window_peak_counts")
    return out


_FACT = [1, 1, 2, 6, 24, 120, 720, 5040]


def _perm_entropy_cpu(w: torch.Tensor) -> torch.Tensor:
    """The outputs of bh_window_perm_entropy for a CPU window (N, B, S)."""
    x = w.double()
    N, B, S = x.shape
    out = torch.zeros((5, B, S), dtype=torch.float64)
    for k, d in enumerate(range(3, 8)):
        npat = N - d + 1
        if npat <= 0:
            continue
        emb = x.unfold(0, d, 1)  # (npat, B, S, d)
        order = torch.argsort(emb, dim=-1, stable=True)
        mult = torch.tensor([_FACT[d - 1 - i] for i in range(d)], dtype=torch.long)
        ids = (order * mult).sum(-1).reshape(npat, B * S)
        cnt = torch.zeros((B * S, _FACT[d]), dtype=torch.float64)
        cnt.scatter_add_(1, ids.T, torch.ones_like(ids.T, dtype=torch.float64))
        q = cnt / npat
        e = -(q * torch.where(q > 0, q.log(), torch.zeros_like(q))).sum(1)
        out[k] = (e / math.log(_FACT[d])).reshape(B, S)
    ok = torch.isfinite(x).all(0)
    return torch.where(ok, out, torch.full_like(out, float("nan")))


def window_perm_entropy(y: torch.Tensor, start: int) -> torch.Tensor:
    """(5, B, S) float64 permutation entropies (tau 1, dimensions 3..7) of the
    window rows [start, N) of y (bh_window_perm_entropy on the device)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        return _perm_entropy_cpu(y[start:])
    y = y.contiguous()
    out = torch.empty((5, B, S), dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_perm_entropy(dt, y.data_ptr(), N, B, S, int(start),
                                                     out.data_ptr(), s.cuda_stream),
                   "bh_window_perm_entropy")
    return out


def peak_feature(name: str, params: dict | None, counts: dict, perm: torch.Tensor | None,
                 n_win: int) -> torch.Tensor:
    """number_peaks (wide supports), number_cwt_peaks and
    permutation_entropy (B, S) float64 (torch_features_pattern.py:64-190,
    torch_features_entropy_complexity.py:11-53).  counts: support -> its
    row of strict counts, and "widths" -> the (8, B, S) per-width counts."""
    p = params or {}
    if name == "number_peaks":
        return counts[int(p["n"])]
    if name == "number_cwt_peaks":
        valid = [wd for wd in range(1, int(p.get("max_width", 5)) + 1) if 2 * wd < n_win]
        cw = counts["widths"]
        if not valid:
            return torch.where(torch.isnan(cw[0]), cw[0], torch.zeros_like(cw[0]))
        return sum(cw[wd - 1] for wd in valid) / len(valid)
    if name == "permutation_entropy":
        return perm[int(p["dimension"]) - 3]
    raise KeyError(f"feature {name!r} is not a peak / pattern feature")


def _corr_dim_cpu(x: torch.Tensor) -> float:
    """correlation_dimension of one float64 series (emb 4, lag 1, 50 radii)."""
    n = x.shape[0]
    m = n - 3
    if m < 2 or not bool(torch.isfinite(x).all()):
        return math.nan
    orb = torch.stack([x[e: e + m] for e in range(4)], 1)
    d = torch.cdist(orb, orb, compute_mode="donot_use_mm_for_euclid_dist")
    d.fill_diagonal_(math.inf)
    sd = float(x.std())
    L0, L1 = math.log(0.1 * sd + 1e-10), math.log(0.5 * sd + 1e-10)
    lr = torch.tensor([L0 + k / 49 * (L1 - L0) for k in range(50)], dtype=torch.float64)
    c = torch.stack([(d <= math.exp(v)).sum() for v in lr.tolist()]).double() / (m * (m - 1))
    ok = c > 0
    nv = int(ok.sum())
    lc = torch.log(c[ok] + 1e-30)
    r = lr[ok]
    den = nv * float((r * r).sum()) - float(r.sum()) ** 2
    if nv < 2 or abs(den) <= 1e-10:
        return math.nan
    return (nv * float((r * lc).sum()) - float(r.sum()) * float(lc.sum())) / den


def _lyap_r_cpu(x: torch.Tensor) -> float:
    """The Rosenstein slope of one float64 series (emb 10, lag 1, 20 steps)."""
    n = x.shape[0]
    m = n - 9
    if m < 1 or not bool(torch.isfinite(x).all()):
        return math.nan
    M = 2 * n - 1
    P = torch.fft.rfft(x, n=M).abs() ** 2
    f = torch.arange(P.shape[0], dtype=torch.float64) / M
    mf = float((f[1:] * P[1:]).sum() / (P[1:].sum() + 1e-10))
    tsep = min(max(math.ceil(1.0 / (mf + 1e-10)), 1), int(0.25 * n))
    orb = torch.stack([x[e: e + m] for e in range(10)], 1)
    d = torch.cdist(orb, orb, compute_mode="donot_use_mm_for_euclid_dist")
    idx = torch.arange(m)
    d = torch.where((idx.view(-1, 1) - idx.view(1, -1)).abs() <= tsep, math.inf, d)
    ntraj = max(m - 19, 1)
    nb = torch.argmin(d[:ntraj, :ntraj], dim=1)
    div = []
    for k in range(20):
        ia, ib = torch.arange(ntraj) + k, nb + k
        ok = (ia < m) & (ib < m)
        v = d[ia[ok], ib[ok]]
        v = v[v > 0]
        div.append(float(torch.log(v).mean()) if v.numel() else math.nan)
    ks = [k for k in range(20) if math.isfinite(div[k])]
    if len(ks) < 2:
        return math.nan
    kk = torch.tensor(ks, dtype=torch.float64)
    dd = torch.tensor([div[k] for k in ks], dtype=torch.float64)
    nf = len(ks)
    den = nf * float((kk * kk).sum()) - float(kk.sum()) ** 2
    if abs(den) <= 1e-10:
        return math.nan
    return (nf * float((kk * dd).sum()) - float(kk.sum()) * float(dd.sum())) / den


def _lyap_e_cpu(x: torch.Tensor) -> list:
    """Eckmann's four exponents of one float64 series (emb 10, matrix 4,
    min_nb 8, min_tsep 0; not divided by tau)."""
    n = x.shape[0]
    if not bool(torch.isfinite(x).all()):
        return [math.inf] * 4
    no = n - 3 - 9
    q = torch.eye(4, dtype=torch.float64)
    lexp = torch.zeros(4, dtype=torch.float64)
    cnt = torch.zeros(4, dtype=torch.float64)
    if no >= 1:
        orb = torch.stack([x[e: e + no] for e in range(10)], 1)
        stp = torch.stack([x[3 * k: 3 * k + no] for k in range(4)], 1)
        beta = x[12: 12 + no]
        for i in range(no):
            d = (orb - orb[i]).abs().max(1).values
            d[i] = math.inf
            r = float(torch.sort(d).values[7]) if no > 8 else math.inf
            if math.isinf(r):
                continue
            nbm = d <= r
            xd = stp[nbm] - stp[i]
            bd = beta[nbm] - beta[i]
            a = torch.linalg.solve(xd.T @ xd + 1e-10 * torch.eye(4, dtype=torch.float64), xd.T @ bd)
            t = torch.zeros((4, 4), dtype=torch.float64)
            t[:3, 1:] = torch.eye(3, dtype=torch.float64)
            t[3] = a
            mq, mr = torch.linalg.qr(t @ q)
            sg = torch.sign(torch.diag(mr))
            sg = torch.where(sg == 0, torch.ones_like(sg), sg)
            q = mq * sg
            dg = torch.diag(mr) * sg
            pos = dg > 0
            lexp[pos] += torch.log(dg[pos])
            cnt[pos] += 1
    return [float(lexp[k] / cnt[k] / 3) if cnt[k] > 0 else math.inf for k in range(4)]


def window_embedding(y: torch.Tensor, start: int, which: str) -> torch.Tensor:
    """Delay-embedding estimates of the window rows [start, N) of y: (B, S)
    float64 correlation_dimension ("corr_dim") or Rosenstein slope
    ("lyap_r"), or (4, B, S) Eckmann exponents ("lyap_e"), the Lyapunov
    estimates not divided by tau (bh_window_embedding on the device, torch
    per series on the CPU)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        w = y[start:].double()
        if which == "lyap_e":
            v = torch.tensor([[_lyap_e_cpu(w[:, b, s]) for s in range(S)] for b in range(B)],
                             dtype=torch.float64)  # (B, S, 4)
            return v.permute(2, 0, 1).contiguous()
        fn = _corr_dim_cpu if which == "corr_dim" else _lyap_r_cpu
        return torch.tensor([[fn(w[:, b, s]) for s in range(S)] for b in range(B)],
                            dtype=torch.float64).reshape(B, S)
    y = y.contiguous()
    shape = (4, B, S) if which == "lyap_e" else (B, S)
    out = torch.empty(shape, dtype=torch.float64, device=y.device)
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    code = {"corr_dim": _lib.BH_EMB_CORR_DIM, "lyap_r": _lib.BH_EMB_LYAP_R,
            "lyap_e": _lib.BH_EMB_LYAP_E}[which]
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_embedding(dt, y.data_ptr(), N, B, S, int(start), code,
                                                  out.data_ptr(), s.cuda_stream),
                   "bh_window_embedding")
    return out


def _lempel_ziv_cpu(w: torch.Tensor, bins: list) -> torch.Tensor:
    """The outputs of bh_window_lempel_ziv for a CPU window (N, B, S)."""
    N, B, S = w.shape
    mn, mx = w.min(0).values, w.max(0).values
    rng = mx - mn + torch.tensor(1e-10, dtype=w.dtype)
    ok = torch.isfinite(w).all(0)
    bn = N / (math.log(N) + 1e-10) if N > 1 else 1.0
    out = torch.empty((len(bins), B, S), dtype=torch.float64)
    for k, nb in enumerate(bins):
        sym = ((w - mn) / rng * nb).long().clamp(0, nb - 1)
        for b in range(B):
            for s in range(S):
                if not bool(ok[b, s]):
                    out[k, b, s] = math.nan
                    continue
                seq = sym[:, b, s].tolist()
                trie: dict = {}
                nodes, count, i = 1, 0, 0
                while i < N:
                    cur, ln, added = 0, 1, False
                    while i + ln <= N:
                        child = trie.get((cur, seq[i + ln - 1]))
                        if child is None:
                            trie[(cur, seq[i + ln - 1])] = nodes
                            nodes += 1
                            count += 1
                            added = True
                            break
                        cur = child
                        ln += 1
                    i += ln
                    if not added:
                        break
                out[k, b, s] = count / bn
    return out


def window_lempel_ziv(y: torch.Tensor, start: int, bins: list) -> torch.Tensor:
    """(len(bins), B, S) float64 lempel_ziv_complexity of the window rows
    [start, N) of y for each bins setting (bh_window_lempel_ziv on the
    device)."""
    N, B, S = y.shape
    if not 0 <= start < N:
        raise ValueError("empty feature window")
    if y.device.type != "cuda":
        return _lempel_ziv_cpu(y[start:], list(bins))
    import ctypes

    y = y.contiguous()
    out = torch.empty((len(bins), B, S), dtype=torch.float64, device=y.device)
    arr = (ctypes.c_int32 * len(bins))(*[int(b) for b in bins])
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    with torch.cuda.device(y.device):
        _lib.check(_lib.lib().bh_window_lempel_ziv(dt, y.data_ptr(), N, B, S, int(start),
                                                   ctypes.cast(arr, ctypes.c_void_p), len(bins),
                                                   out.data_ptr(), s.cuda_stream),
                   "bh_window_lempel_ziv")
    return out


def feature_family(name: str, params: dict | None = None) -> str:
    """Which kernel output a comprehensive feature is formed from: "block"
    (bh_window_features), "spectrum", "sorted", "autocov", "trends" or
    "peaks" (number_peaks of supports other than 1, 3, 5 among them)."""
    if name == "number_peaks" and params and int(params.get("n", 1)) not in _BLOCK_PEAKS:
        return "peaks"
    if name in PEAK_FEATURES:
        return "peaks"
    if name in EMBED_FEATURES:
        return "embedding"
    if name in LZ_FEATURES:
        return "lempel_ziv"
    if name in SPECTRUM_FEATURES:
        return "spectrum"
    if name in SORTED_FEATURES:
        return "sorted"
    if name in AUTOCOV_FEATURES:
        return "autocov"
    if name in TRENDS_FEATURES:
        return "trends"
    if name in COMPREHENSIVE_FEATURES:
        return "block"
    raise KeyError(f"feature {name!r} is not a comprehensive window feature")


class WindowBlocks:
    """The kernel outputs of one kept window (rows [start, N) of y), each
    computed once, when a feature of its family is first asked for: the
    window-feature slots, the spectrum (as many bins as the plan's spectrum
    features need), the sorted-window statistics, the autocovariances (to
    the largest lag the plan reads), the chunk trends and fits.  median: a callable giving
    the window median (for symmetry_looking)."""

    def __init__(self, y: torch.Tensor, start: int, plan, median=None):
        self.y, self.start, self.n = y, int(start), y.shape[0] - int(start)
        self.plan = list(plan)
        self._median_fn = median
        self._out: dict = {}
        self.memo: dict = {}

    def _get(self, key, make):
        if key not in self._out:
            self._out[key] = make()
        return self._out[key]

    def block(self) -> torch.Tensor:
        return self._get("block", lambda: window_feature_block(self.y, self.start))

    def spectrum(self):
        return self._get("spectrum", lambda: window_spectrum(self.y, self.start,
                                                             *spectrum_needs(self.plan)))

    def sorted(self):
        cq = any(name == "change_quantiles" for _, name, _ in self.plan)
        return self._get("sorted", lambda: window_sorted(self.y, self.start, cq))

    def autocov(self):
        return self._get("autocov", lambda: window_autocov(self.y, self.start,
                                                           autocov_needs(self.plan)))

    def trends(self):
        cwt = any(name == "cwt_coefficients" for _, name, _ in self.plan)
        return self._get("trends", lambda: window_trends(self.y, self.start, cwt))

    def peak_counts(self) -> dict:
        def make():
            sup = sorted({int(p["n"]) for _, nm, p in self.plan
                          if nm == "number_peaks" and p and int(p["n"]) not in _BLOCK_PEAKS})
            rows: dict = {}
            for i in range(0, max(len(sup), 1), 4):
                chunk = sup[i:i + 4]
                out = window_peak_counts(self.y, self.start, chunk)
                rows.update({k: out[r] for r, k in enumerate(chunk)})
                rows["widths"] = out[4:]
            return rows
        return self._get("peaks", make)

    def perm_entropy(self):
        return self._get("perm", lambda: window_perm_entropy(self.y, self.start))

    def lempel_ziv(self) -> dict:
        def make():
            bins = sorted({int(p["bins"]) for _, nm, p in self.plan if nm == "lempel_ziv_complexity"})
            rows: dict = {}
            for i in range(0, len(bins), 8):
                chunk = bins[i:i + 8]
                out = window_lempel_ziv(self.y, self.start, chunk)
                rows.update({b: out[k] for k, b in enumerate(chunk)})
            return rows
        return self._get("lz", make)

    def feature(self, name: str, params: dict | None) -> torch.Tensor:
        """Comprehensive feature `name` (B, S) float64 of the window."""
        if name == "lempel_ziv_complexity":  # entropy_complexity.py:140-180
            return self.lempel_ziv()[int((params or {})["bins"])]
        if name == "correlation_dimension":
            return self._get("corr_dim", lambda: window_embedding(self.y, self.start, "corr_dim"))
        if name == "lyapunov_r":  # torch_features_dynamical.py:23-42: the slope over tau
            slope = self._get("lyap_r", lambda: window_embedding(self.y, self.start, "lyap_r"))
            return slope / float((params or {}).get("tau", 1.0))
        if name == "lyapunov_e":  # torch_features_dynamical.py:45-66: (4, B, S), over tau
            ex = self._get("lyap_e", lambda: window_embedding(self.y, self.start, "lyap_e"))
            return ex / float((params or {}).get("tau", 1.0))
        if feature_family(name, params) == "peaks":
            perm = self.perm_entropy() if name == "permutation_entropy" else None
            counts = self.peak_counts() if name != "permutation_entropy" else {}
            return peak_feature(name, params, counts, perm, self.n)
        if name in TRENDS_FEATURES:
            return trends_feature(name, params, self.trends())
        if name in AUTOCOV_FEATURES:
            return autocov_feature(name, params, self.autocov(), self.n, self.memo)
        if name in SPECTRUM_FEATURES:
            return spectrum_feature(name, params, *self.spectrum(), self.n)
        if name in SORTED_FEATURES:
            return sorted_feature(name, params, *self.sorted()[:2], self.n)
        med = None
        if name == "symmetry_looking":
            if self.n <= (8192 if self.y.dtype == torch.float32 else 4096) or self._median_fn is None:
                med = self.sorted()[2]  # the sorted window's middle value
            else:
                med = self._get("median", self._median_fn)
        return comprehensive_feature(name, params, self.block(), self.n, med, self.memo)
