"""Steady-state features: filter_time, the streaming statistical set, fused extraction.

* ``Solution``            data contract of src/pybasin/solution.py:24-91
                          (y is (N, B, S), N == len(t)).
* ``FeatureExtractor``    base with ``filter_time``
                          (src/pybasin/feature_extractors/feature_extractor.py:85-130).
* ``StatFeatureExtractor`` the streaming subset of TorchFeatureExtractor
                          (src/pybasin/ts_torch/torch_feature_extractor.py:192-285):
                          per-state feature dicts, columns ``state_{i}__{name}``
                          state-major then dict order, impute_extreme (+-inf/NaN
                          -> +-1e10, torch_feature_utilities.py:60-94) and an
                          optional z-score fitted on the first call.
* ``HipFusedFeatureExtractor`` the same columns taken from the statistics the
                          HIP kernel accumulated during integration; falls back
                          to the trajectory when the solution carries none
                          (e.g. a CPU template clone).
"""

from __future__ import annotations

import threading
from typing import Any

import torch

from .solvers.hip_dopri5 import FUSED_RESULTS, FusedResult, steady_index
from .window_features import COMPREHENSIVE_FEATURES, MULTI_OUTPUT, WindowBlocks, check_params

DEFAULT_STEADY_FRACTION = 0.85

# every feature here is computable from mean / std / min / max
STREAMING_FEATURES = ("mean", "standard_deviation", "variance", "root_mean_square", "minimum",
                      "maximum", "absolute_maximum", "delta", "log_delta", "amplitude")
# the least-squares slope of the window, a fifth streaming statistic of the
# kernel (integrate with trend=True); attr "slope" only
TREND_FEATURES = ("linear_trend",)
# features of the whole window series (need the saved window: keep_trajectories):
# order statistics, autocorrelation and spectrum
WINDOW_FEATURES = ("median", "quantile", "autocorrelation_periodicity", "spectral_frequency_ratio")
DEFAULT_FEATURES = {"mean": None, "standard_deviation": None}
# the calculators of the reference's comprehensive set reduced by
# bh_window_features (window_features.py); they need the saved window too


def is_comprehensive(name: str, params: dict | None) -> bool:
    """Whether a requested feature is formed from the window kernel's slots
    (linear_trend with an attribute other than the fused slope is)."""
    if name in STREAMING_FEATURES or name in WINDOW_FEATURES:
        return False
    if name == "linear_trend":
        return (params or {}).get("attr", "slope") != "slope"
    return name in COMPREHENSIVE_FEATURES
# the reference's default set (settings.py:376-395: minimal features + delta, log_delta)
MINIMAL_FEATURES = dict.fromkeys(["median", "mean", "standard_deviation", "variance",
                                  "root_mean_square", "maximum", "absolute_maximum", "minimum",
                                  "delta", "log_delta"])
# the reference's dynamical-system set (settings.py:407-417)
DYNAMICAL_SYSTEM_FEATURES = {
    "variance": None,
    "amplitude": None,
    "mean": None,
    "linear_trend": [{"attr": "slope"}],
    "autocorrelation_periodicity": [{"output": "strength"}, {"output": "period"}],
    "spectral_frequency_ratio": None,
}


class Solution:
    def __init__(self, initial_condition: torch.Tensor, time: torch.Tensor, y: torch.Tensor,
                 features: torch.Tensor | None = None, labels=None, model_params=None):
        assert initial_condition.ndim == 2, "initial_condition must be 2D (B, S)"
        assert time.ndim == 1, "time must be 1D (N,)"
        assert y.ndim == 3, "y must be 3D (N, B, S)"
        assert y.shape[0] == time.shape[0], "time dimension mismatch"
        assert y.shape[1] == initial_condition.shape[0], "initial conditions dimension mismatch"
        assert y.shape[2] == initial_condition.shape[1], "state dimension mismatch"
        self.initial_condition = initial_condition
        self.time = time
        self.y = y
        self.features = features
        self.extracted_features = None
        self.extracted_feature_names = None
        self.feature_names = None
        self.labels = labels
        self.model_params = model_params


class FeatureExtractor:
    def __init__(self, time_steady: float | None = None):
        self.time_steady = time_steady

    def _resolve_time_steady(self, solution: Solution) -> float:
        if self.time_steady is not None:
            return self.time_steady
        t0, t1 = float(solution.time[0]), float(solution.time[-1])
        return t0 + DEFAULT_STEADY_FRACTION * (t1 - t0)

    def filter_time(self, solution: Solution) -> torch.Tensor:
        ts = self._resolve_time_steady(solution)
        idx = torch.where(solution.time >= ts)[0]
        start = int(idx[0]) if len(idx) > 0 else 0
        return solution.y[start:, ...]

    def extract_features(self, solution: Solution) -> torch.Tensor:  # pragma: no cover
        raise NotImplementedError


def impute_extreme(features: torch.Tensor, extreme_value: float = 1e10) -> torch.Tensor:
    out = features.clone()
    ext = torch.tensor(extreme_value, dtype=features.dtype, device=features.device)
    out = torch.where(torch.isposinf(features), ext, out)
    out = torch.where(torch.isneginf(features), -ext, out)
    out = torch.where(torch.isnan(features), ext, out)
    return out


def stats_from_window(window: torch.Tensor, trend: bool = False) -> torch.Tensor:
    """(N', B, S) -> (B, S, 4) float64: mean, std (ddof 0), min, max; with
    trend=True (B, S, 5), the least-squares slope against the sample
    position appended (torch_features_trend.py:11-27)."""
    w = window.to(torch.float64)
    cols = [w.mean(0), w.std(0, correction=0), w.min(0).values, w.max(0).values]
    if trend:
        n = w.shape[0]
        t = torch.arange(n, dtype=torch.float64, device=w.device)
        tc = (t - t.mean()).reshape(n, 1, 1)
        ss_tt = float(((t - t.mean()) ** 2).sum())
        cols.append((tc * (w - cols[0])).sum(0) / (ss_tt + 1e-10))
    return torch.stack(cols, -1)


def feature_from_stats(name: str, stats: torch.Tensor) -> torch.Tensor:
    """One streaming feature (torch_features_statistical.py:39-110) from (B, S, 4) stats."""
    mean, std, mn, mx = stats.unbind(-1)[:4]
    if name == "mean":
        return mean
    if name == "standard_deviation":
        return std
    if name == "variance":
        return std * std
    if name == "root_mean_square":
        return torch.sqrt(std * std + mean * mean)
    if name == "minimum":
        return mn
    if name == "maximum":
        return mx
    if name == "absolute_maximum":
        return torch.maximum(mx.abs(), mn.abs())
    if name == "delta":
        return (mx - mean).abs()
    if name == "log_delta":
        return torch.log((mx - mean).abs() + 1e-10)
    if name == "amplitude":
        return mx - mn
    raise KeyError(f"feature {name!r} is not a streaming feature; available: {STREAMING_FEATURES}")


def feature_name(state: int, name: str, params: dict | None) -> str:
    """state_{i}__{name}[__{key}_{value} ...] with parameters sorted by key
    (src/pybasin/feature_extractors/utils.py:34-80)."""
    out = f"state_{state}__{name}"
    for k in sorted(params or {}):
        out += f"__{k}_{params[k]}"
    return out


def window_order_stats(y: torch.Tensor, start: int, ranks: list[int]) -> torch.Tensor:
    """(n_ranks, B, S): the values of the given 0-based ranks among rows
    [start, N) of a (N, B, S) trajectory tensor; NaN where the window holds
    a NaN.  HIP kernel (bh_window_order_stats) for device tensors, torch sort
    on the CPU."""
    if y.device.type == "cuda":
        from . import _lib

        y = y.contiguous()
        N, B, S = y.shape
        r = torch.tensor(ranks, dtype=torch.int64, device=y.device)
        out = torch.empty((len(ranks), B, S), dtype=y.dtype, device=y.device)
        dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
        s = torch.cuda.current_stream(y.device)
        with torch.cuda.device(y.device):
            _lib.check(_lib.lib().bh_window_order_stats(dt, y.data_ptr(), N, B, S, int(start),
                                                         r.data_ptr(), len(ranks), out.data_ptr(),
                                                         s.cuda_stream), "bh_window_order_stats")
        return out
    w = y[start:]
    srt = torch.sort(w, dim=0).values
    out = srt[ranks]
    return torch.where(torch.isnan(w).any(0, keepdim=True), torch.full_like(out, float("nan")), out)


def _local_maxima_first(ac: torch.Tensor, min_lag: int, thr: float):
    """First local maximum (scipy's definition, plateaus at their midpoint;
    torch_feature_utilities.py:146-211) of each column of ac (N, M) at an
    index >= min_lag with a value > thr: (values, indices), 0 where none."""
    n, m = ac.shape
    val = torch.zeros(m, dtype=ac.dtype)
    idx = torch.zeros(m, dtype=ac.dtype)
    a = ac.cpu().tolist()
    for j in range(m):
        col = [row[j] for row in a]
        i = 1
        while i < n - 1:
            if col[i - 1] < col[i]:
                k = i + 1
                while k < n - 1 and col[k] == col[i]:
                    k += 1
                if col[k] < col[i]:
                    mid = (i + k - 1) // 2
                    if mid >= min_lag and col[mid] > thr:
                        val[j], idx[j] = col[mid], float(mid)
                        break
                    i = k
            i += 1
    return val, idx


def _series_cpu(name: str, params: dict | None, w: torch.Tensor) -> torch.Tensor:
    """Torch restatement of the series features on a CPU window (N', B, S)
    (the CPU template path): the reference's FFT formulation."""
    n, B, S = w.shape
    x = w.to(torch.float64)
    if name == "autocorrelation_periodicity":
        p = params or {}
        z = (x - x.mean(0, keepdim=True)) / (x.std(0, keepdim=True) + 1e-10)
        f = torch.fft.rfft(z, n=2 * n, dim=0)
        ac = torch.fft.irfft(f.real ** 2 + f.imag ** 2, n=2 * n, dim=0)[:n]
        ac = ac / (ac[0:1] + 1e-10)
        val, idx = _local_maxima_first(ac.reshape(n, -1), int(p.get("min_lag", 2)),
                                       float(p.get("peak_threshold", 0.3)))
        out = idx if p.get("output", "strength") == "period" else val
        return out.reshape(B, S).to(w.dtype)
    c = x - x.mean(0, keepdim=True)
    f = torch.fft.rfft(c, dim=0)
    pw = (f.real ** 2 + f.imag ** 2).reshape(f.shape[0], -1)
    pw[0] = 0.0
    out = torch.zeros(pw.shape[1], dtype=torch.float64)
    bins = torch.arange(pw.shape[0])
    for j in range(pw.shape[1]):
        psd = pw[:, j]
        if not float(psd.max()) >= 1e-10:
            continue
        first = int(torch.argmax(psd))
        cand = torch.where((bins - first).abs() >= 2, psd, torch.full_like(psd, -1.0))
        second = int(torch.argmax(cand))
        if float(cand[second]) > 0 and first > 0:
            out[j] = second / first
    return out.reshape(B, S).to(w.dtype)


def _series_hip(name: str, params: dict | None, y: torch.Tensor, start: int) -> torch.Tensor:
    from . import _lib

    y = y.contiguous()
    N, B, S = y.shape
    dt = _lib.BH_F32 if y.dtype == torch.float32 else _lib.BH_F64
    s = torch.cuda.current_stream(y.device)
    out = torch.empty((B, S), dtype=y.dtype, device=y.device)
    with torch.cuda.device(y.device):
        if name == "autocorrelation_periodicity":
            p = params or {}
            other = torch.empty_like(out)
            strength, period = (out, other) if p.get("output", "strength") != "period" else (other, out)
            _lib.check(_lib.lib().bh_window_autocorr_periodicity(
                dt, y.data_ptr(), N, B, S, int(start), int(p.get("min_lag", 2)),
                float(p.get("peak_threshold", 0.3)), strength.data_ptr(), period.data_ptr(),
                s.cuda_stream), "bh_window_autocorr_periodicity")
        else:
            _lib.check(_lib.lib().bh_window_spectral_ratio(dt, y.data_ptr(), N, B, S, int(start),
                                                            out.data_ptr(), s.cuda_stream),
                       "bh_window_spectral_ratio")
    return out


def window_feature(name: str, params: dict | None, y: torch.Tensor, start: int) -> torch.Tensor:
    """A feature of the window rows [start, N) of y (N, B, S) -> (B, S):
    median / quantile as x.median(dim=0).values and torch.quantile(x, q, dim=0)
    (torch_features_statistical.py:30-38, 154-157); autocorrelation_periodicity
    (output strength or period; torch_features_autocorrelation.py:268-331) and
    spectral_frequency_ratio (torch_features_frequency.py:390-441) by the HIP
    series kernels on the device, by a torch restatement on the CPU."""
    n = y.shape[0] - start
    if n < 1:
        raise ValueError("empty feature window")
    if name in ("autocorrelation_periodicity", "spectral_frequency_ratio"):
        if y.device.type == "cuda":
            return _series_hip(name, params, y, start)
        return _series_cpu(name, params, y[start:])
    if name == "median":
        return window_order_stats(y, start, [(n - 1) // 2])[0]
    if name == "quantile":
        q = float(params["q"])
        pos = q * (n - 1)
        lo, hi = int(torch.floor(torch.tensor(pos))), int(torch.ceil(torch.tensor(pos)))
        v = window_order_stats(y, start, [lo, hi])
        return torch.lerp(v[0], v[1], torch.tensor(pos - lo, dtype=v.dtype, device=v.device))
    raise KeyError(f"feature {name!r} is not a window feature; available: {WINDOW_FEATURES}")


class StatFeatureExtractor(FeatureExtractor):
    """Streaming statistical features, column-compatible with TorchFeatureExtractor."""

    def __init__(self, time_steady: float | None = None,
                 features: dict[str, Any] | None = DEFAULT_FEATURES,  # type: ignore[assignment]
                 features_per_state: dict[int, dict[str, Any] | None] | None = None,
                 normalize: bool = False):
        super().__init__(time_steady)
        self.features = features
        self.features_per_state = features_per_state or {}
        self.normalize = normalize
        self._mean: torch.Tensor | None = None
        self._std: torch.Tensor | None = None
        self._lock = threading.Lock()
        self._names: list[str] = []

    def _plan(self, n_states: int) -> list[tuple[int, str, dict | None]]:
        plan = []
        for s in range(n_states):
            if s in self.features_per_state:
                fdict = self.features_per_state[s]
                if fdict is None:
                    continue
            else:
                fdict = self.features
                if fdict is None:
                    continue
            for name, plist in fdict.items():
                known = STREAMING_FEATURES + TREND_FEATURES + WINDOW_FEATURES
                if name not in known and name not in COMPREHENSIVE_FEATURES:
                    raise KeyError(f"feature {name!r} is not supported; available: "
                                   f"{sorted(set(known) | set(COMPREHENSIVE_FEATURES))}")
                for params in (plist if plist else [None]):
                    if is_comprehensive(name, params):
                        check_params(name, params)
                    plan.append((s, name, params))
        return plan

    def _columns(self, plan, stats: torch.Tensor, y: torch.Tensor | None,
                 start: int) -> torch.Tensor:
        cols = []
        names = []
        wb = None
        for s, name, params in plan:
            base = feature_name(s, name, params)
            if is_comprehensive(name, params):
                if y is None:
                    raise ValueError(f"{name} needs the saved window: integrate with "
                                     "keep_trajectories=True")
                if wb is None:
                    wb = WindowBlocks(y, start, plan,
                                      median=lambda: window_feature("median", None, y, start))
                v = wb.feature(name, params)
                if name in MULTI_OUTPUT:  # one column per output, named __0, __1, ..
                    for k in range(MULTI_OUTPUT[name]):
                        cols.append(v[k][:, s].to(stats.device, torch.float64))
                        names.append(f"{base}__{k}")
                    continue
                cols.append(v[:, s].to(stats.device, torch.float64))
            elif name in WINDOW_FEATURES:
                if y is None:
                    raise ValueError(f"{name} needs the saved window: integrate with "
                                     "keep_trajectories=True")
                cols.append(window_feature(name, params, y, start)[:, s].to(stats.device,
                                                                             torch.float64))
            elif name in TREND_FEATURES:
                if stats.shape[-1] >= 5:
                    cols.append(stats[:, s, 4])
                elif y is not None:
                    cols.append(stats_from_window(y[start:], trend=True)[:, s, 4].to(stats.device))
                else:
                    raise ValueError("linear_trend needs the slope statistic: integrate with "
                                     "trend=True (or keep_trajectories=True)")
            else:
                cols.append(feature_from_stats(name, stats)[:, s])
            names.append(base)
        self._names = names
        feats = torch.stack(cols, dim=1) if cols else stats.new_zeros((stats.shape[0], 0))
        return impute_extreme(feats.to(torch.float32))

    @property
    def feature_names(self) -> list[str]:
        return list(self._names)

    def _stats(self, solution: Solution) -> torch.Tensor:
        return stats_from_window(self.filter_time(solution), trend=self.needs_trend)

    def _requested(self) -> set:
        dicts = [self.features] + list(self.features_per_state.values())
        return {name for d in dicts if d for name in d}

    @property
    def needs_window(self) -> bool:
        """Whether some requested feature reads the whole window series
        (order statistics, autocorrelation, spectrum: the saved window must
        be kept)."""
        dicts = [self.features] + list(self.features_per_state.values())
        return any(name in WINDOW_FEATURES or is_comprehensive(name, params)
                   for d in dicts if d for name, plist in d.items()
                   for params in (plist if plist else [None]))

    @property
    def needs_trend(self) -> bool:
        """Whether the slope statistic is requested (integrate with trend=True)."""
        return any(name in TREND_FEATURES for name in self._requested())

    def _normalized(self, feats: torch.Tensor) -> torch.Tensor:
        # z-score with the mean / std of the first batch seen (then fixed)
        if not self.normalize:
            return feats
        with self._lock:
            if self._mean is None:
                self._mean = feats.mean(0, keepdim=True)
                std = feats.std(0, keepdim=True)
                # a constant column keeps its scale (torch_feature_extractor.py:273-281)
                self._std = torch.where(std == 0, torch.ones_like(std), std)
        return (feats - self._mean) / self._std

    def extract_features(self, solution: Solution) -> torch.Tensor:
        stats = self._stats(solution)
        plan = self._plan(stats.shape[1])
        start = steady_index(solution.time, self._resolve_time_steady(solution))
        return self._normalized(self._columns(plan, stats, solution.y, start))


class HipFusedFeatureExtractor(StatFeatureExtractor):
    """StatFeatureExtractor fed by the kernel's fused window statistics."""

    def _stats(self, solution: Solution) -> torch.Tensor:
        res: FusedResult | None = FUSED_RESULTS.get(solution.y)
        if res is not None and res.stats is not None:
            want = steady_index(solution.time, self._resolve_time_steady(solution))
            if want == res.steady_idx and (res.stats.shape[-1] >= 5 or not self.needs_trend):
                return res.stats
            if res.y is not None:
                return stats_from_window(res.y[want:], trend=self.needs_trend)
        return super()._stats(solution)

    def extract_from_result(self, res: FusedResult) -> torch.Tensor:
        """Features straight from a fused result (no Solution / trajectory)."""
        if res.stats is None:
            raise ValueError("the fused result carries no window statistics")
        plan = self._plan(res.stats.shape[1])
        return self._normalized(self._columns(plan, res.stats, res.y, res.steady_idx))
