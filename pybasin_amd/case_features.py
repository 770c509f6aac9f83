"""The case studies' own feature extractors, fed by the fused window statistics.

Each extractor reads only per-state mean / std / min / max of the steady
window, so the HIP kernel's fused statistics (``FusedResult.stats``) replace
the trajectory; with a plain ``Solution`` the same statistics are computed
from the window.  Outputs follow the reference classes column for column:

* ``PendulumFeatureExtractor``  case_studies/pendulum/pendulum_feature_extractor.py:12-29
  one-hot float64 [FP, LC]; FP iff |max(omega) - mean(omega)| < 0.01.
* ``DuffingFeatureExtractor``   case_studies/duffing_oscillator/duffing_feature_extractor.py:7-37
  [max x, std x] with torch.std's default correction=1 (ddof 1), in the
  state dtype.
* ``LorenzFeatureExtractor``    case_studies/lorenz/lorenz_feature_extractor.py:11-43
  one-hot float64 [S1, S2] by the sign of mean x; [0, 0] when any state
  reaches |y| > 195 in the window (unbounded).

The window statistics are accumulated in float64 where the reference reduces
in the state dtype; the decisions differ only for windows within rounding of
a threshold.
"""

from __future__ import annotations

import math

import torch

from .features import FeatureExtractor, Solution, stats_from_window
from .solvers.hip_dopri5 import FUSED_RESULTS, FusedResult, steady_index


class _WindowStatsExtractor(FeatureExtractor):
    """Base: (B, S, 4) window statistics and the window length from either source."""

    def _window(self, solution: Solution) -> tuple[torch.Tensor, int, torch.dtype]:
        res: FusedResult | None = FUSED_RESULTS.get(solution.y)
        start = steady_index(solution.time, self._resolve_time_steady(solution))
        if res is not None and res.stats is not None and res.steady_idx == start:
            return res.stats, len(solution.time) - start, solution.y.dtype
        w = self.filter_time(solution)
        return stats_from_window(w), w.shape[0], solution.y.dtype

    def _from_stats(self, stats: torch.Tensor, n_window: int, dtype: torch.dtype) -> torch.Tensor:
        raise NotImplementedError

    def extract_features(self, solution: Solution) -> torch.Tensor:
        return self._from_stats(*self._window(solution))

    def extract_from_result(self, res: FusedResult) -> torch.Tensor:
        if res.stats is None:
            raise ValueError("the fused result carries no window statistics")
        return self._from_stats(res.stats, len(res.t) - res.steady_idx, res.final.dtype)


class PendulumFeatureExtractor(_WindowStatsExtractor):
    @property
    def feature_names(self) -> list[str]:
        return ["FP", "LC"]

    def _from_stats(self, stats, n_window, dtype):
        mean, _, _, mx = stats[:, 1].unbind(-1)
        fp = (mx - mean).abs() < 0.01
        out = torch.empty((stats.shape[0], 2), dtype=torch.float64, device=stats.device)
        out[:, 0] = fp.to(torch.float64)
        out[:, 1] = (~fp).to(torch.float64)
        return out.cpu()


class DuffingFeatureExtractor(_WindowStatsExtractor):
    @property
    def feature_names(self) -> list[str]:
        return ["max", "std"]

    def _from_stats(self, stats, n_window, dtype):
        _, std0, _, mx = stats[:, 0].unbind(-1)
        # ddof 1 from ddof 0; a one-point window has no ddof-1 std (torch gives NaN)
        scale = math.sqrt(n_window / (n_window - 1)) if n_window > 1 else float("nan")
        return torch.stack([mx, std0 * scale], dim=1).to(dtype)


class LorenzFeatureExtractor(_WindowStatsExtractor):
    @property
    def feature_names(self) -> list[str]:
        return ["S1", "S2"]

    def _from_stats(self, stats, n_window, dtype):
        mean_x = stats[:, 0, 0]
        absmax = torch.maximum(stats[..., 2].abs(), stats[..., 3].abs())
        unbounded = (absmax > 195).any(dim=1)
        pos = mean_x > 0
        out = torch.zeros((stats.shape[0], 2), dtype=torch.float64, device=stats.device)
        out[:, 0] = (~unbounded & pos).to(torch.float64)
        out[:, 1] = (~unbounded & ~pos).to(torch.float64)
        return out.cpu()
