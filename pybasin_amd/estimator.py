"""A small basin-stability driver around the HIP path.

Restates the pieces of the reference's pipeline that sit next to the solver:
* ``UniformRandomSampler``  src/pybasin/sampler.py:56-65 (float32 limits,
  torch.rand(n, S) * (max - min) + min);
* ``set_seed``              src/pybasin/utils.py:28-41;
* ``KNNTemplateClassifier`` the case studies' KNeighborsClassifier(n_neighbors=1)
  fitted on template-IC features (e.g. case_studies/pendulum/setup_pendulum_system.py:53-65),
  evaluated on the GPU;
* ``compute_bs`` / ``get_errors``  src/pybasin/basin_stability_estimator.py:586-625
  (fractions, e_abs = sqrt(S (1 - S) / n), e_rel = 1 / sqrt(n S));
* ``BasinStabilityEstimator.run``  the integrate -> features -> classify -> BS
  sequence of _run_basin_stability (basin_stability_estimator.py:230-361), with
  the template ICs integrated in the same launch as the sample.
"""

from __future__ import annotations

import random
from dataclasses import dataclass, field

import numpy as np
import torch

from .features import HipFusedFeatureExtractor, Solution, StatFeatureExtractor
from .solvers.hip_dopri5 import FUSED_RESULTS, functor_id


def set_seed(seed: int) -> None:
    random.seed(seed)
    np.random.seed(seed)
    torch.manual_seed(seed)
    if torch.cuda.is_available():
        torch.cuda.manual_seed_all(seed)


class UniformRandomSampler:
    def __init__(self, min_limits, max_limits, device: str | None = None):
        assert len(min_limits) == len(max_limits), "min_limits and max_limits must have the same length"
        if device is None:
            self.device = torch.device("cuda:0" if torch.cuda.is_available() else "cpu")
        else:
            d = torch.device(device)
            self.device = torch.device("cuda:0") if d.type == "cuda" and d.index is None else d
        self.min_limits = torch.tensor(min_limits, dtype=torch.float32, device=self.device)
        self.max_limits = torch.tensor(max_limits, dtype=torch.float32, device=self.device)
        self.state_dim = len(min_limits)

    def sample(self, n: int) -> torch.Tensor:
        return (torch.rand(n, self.state_dim, device=self.device)
                * (self.max_limits - self.min_limits) + self.min_limits)


class KNNTemplateClassifier:
    """1-nearest-neighbour over template features (Euclidean), on any device."""

    def __init__(self, labels: list[str]):
        self.labels = list(labels)
        self._X: torch.Tensor | None = None

    def fit(self, X: torch.Tensor) -> "KNNTemplateClassifier":
        self._X = X.to(torch.float64)
        return self

    def predict_index(self, X: torch.Tensor) -> torch.Tensor:
        assert self._X is not None, "fit() first"
        d = torch.cdist(X.to(torch.float64), self._X.to(X.device))
        return torch.argmin(d, dim=1)

    def predict(self, X: torch.Tensor) -> np.ndarray:
        idx = self.predict_index(X).cpu().numpy()
        return np.asarray(self.labels, dtype=object)[idx]


def compute_bs(labels: np.ndarray) -> dict[str, float]:
    labels = np.asarray(labels)
    n = len(labels)
    uniq, counts = np.unique(labels, return_counts=True)
    return {str(u): float(c) / n for u, c in zip(uniq, counts)}


def get_errors(bs: dict[str, float], n: int) -> dict[str, dict[str, float]]:
    out = {}
    for k, s in bs.items():
        e_abs = float(np.sqrt(s * (1 - s) / n))
        e_rel = float(1.0 / np.sqrt(n * s)) if s > 0 else float("inf")
        out[k] = {"e_abs": e_abs, "e_rel": e_rel}
    return out


@dataclass
class BSResult:
    basin_stability: dict[str, float]
    errors: dict[str, dict[str, float]]
    labels: np.ndarray
    features: torch.Tensor
    feature_names: list[str]
    y0: torch.Tensor
    extra: dict = field(default_factory=dict)


class BasinStabilityEstimator:
    """integrate -> fused features -> 1-NN template labels -> basin stability."""

    def __init__(self, ode_system, sampler: UniformRandomSampler, n: int, solver,
                 feature_extractor: StatFeatureExtractor, template_y0, template_labels,
                 keep_trajectories: bool = False):
        self.ode_system = ode_system
        self.sampler = sampler
        self.n = int(n)
        self.solver = solver
        self.feature_extractor = feature_extractor
        self.template_y0 = torch.as_tensor(template_y0, dtype=torch.float32)
        self.classifier = KNNTemplateClassifier(template_labels)
        self.keep_trajectories = keep_trajectories

    def run(self, y0: torch.Tensor | None = None) -> BSResult:
        if y0 is None:
            y0 = self.sampler.sample(self.n)
        n = y0.shape[0]
        dev = self.solver.device
        batch = torch.cat([y0.to(dev), self.template_y0.to(dev, y0.dtype)])
        fx = self.feature_extractor
        if dev.type == "cuda" and functor_id(self.ode_system) is not None:
            save_ts, batch_p = self.solver._prepare_tensors(batch)
            steady = fx.time_steady
            res = self.solver.run_fused(self.ode_system, batch_p, save_ts, None,
                                        keep_trajectories=self.keep_trajectories,
                                        time_steady=steady)
            if isinstance(fx, HipFusedFeatureExtractor):
                feats = fx.extract_from_result(res)
            else:
                if res.y is None:
                    raise ValueError("a non-fused extractor needs keep_trajectories=True")
                FUSED_RESULTS[res.y] = res
                feats = fx.extract_features(Solution(batch, save_ts, res.y))
            extra = {"status": res.status, "nsteps": res.nsteps}
        else:
            t, y = self.solver.integrate(self.ode_system, batch)
            feats = fx.extract_features(Solution(batch, t, y))
            extra = {}
        self.classifier.fit(feats[n:])
        labels = self.classifier.predict(feats[:n])
        bs = compute_bs(labels)
        return BSResult(bs, get_errors(bs, n), labels, feats[:n], fx.feature_names, y0, extra)
