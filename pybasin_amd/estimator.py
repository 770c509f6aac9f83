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

from .features import MINIMAL_FEATURES, HipFusedFeatureExtractor, Solution, StatFeatureExtractor
from . import _lib
from .solvers.hip_dopri5 import FUSED_RESULTS, HipDopri5Solver, functor_id


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
        if X.device.type == "cuda":
            from .cluster import knn  # HIP kernel (bh_knn)

            _, idx = knn(X.to(torch.float64).contiguous(), self._X.to(X.device).contiguous(), 1)
            return idx[:, 0]
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
    study_label: dict = field(default_factory=lambda: {"baseline": True})


def build_params_grid(ode_system, study: list[dict]) -> torch.Tensor:
    """(P, n_params) float64 grid: the system's params in declaration order with
    each study entry's overrides (basin_stability_estimator.py:469-486)."""
    names = list(dict(ode_system.params).keys())
    base = [float(v) for v in dict(ode_system.params).values()]
    rows = []
    for entry in study:
        row = list(base)
        for k, v in entry.items():
            row[names.index(k)] = float(v)
        rows.append(row)
    return torch.tensor(rows, dtype=torch.float64)


class BasinStabilityEstimator:
    """integrate -> fused features -> 1-NN template labels -> basin stability.

    The template ICs are integrated in the same launch as the sample, with the
    system's own parameters (TemplateIntegrator integrates them once with its
    fixed ode_params, template_integrator.py:59-105).  Unbounded trajectories
    (stopped at the divergence threshold with Inf fill, or Inf in a kept
    trajectory: _detect_unbounded_trajectories, basin_stability_estimator.py:519-527)
    are labelled "unbounded" and left out of classification
    (_process_study_label, 365-461).
    """

    def __init__(self, ode_system, sampler: UniformRandomSampler, n: int = 10_000, solver=None,
                 feature_extractor=None, template_y0=None, template_labels=None,
                 keep_trajectories: bool = False, detect_unbounded: bool = True,
                 compute_orbit_data: list[int] | bool = False, predictor=None):
        # defaults as the reference's (basin_stability_estimator.py:73-181):
        # Dopri5 over t in [0, 1000] saving 1,000 points of [850, 1000] on the
        # sampler's device, the minimal feature set z-scored, and without
        # templates HDBSCAN with auto-tuned min_cluster_size and noise assignment
        if solver is None:
            solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(850.0, 1000.0),
                                     device=str(sampler.device))
        if feature_extractor is None:
            feature_extractor = HipFusedFeatureExtractor(features=MINIMAL_FEATURES, normalize=True)
        if getattr(feature_extractor, "needs_window", False):
            keep_trajectories = True  # order statistics read the saved window
        if template_y0 is None and predictor is None:
            from .cluster import HipHDBSCANClusterer

            predictor = HipHDBSCANClusterer(auto_tune=True, assign_noise=True,
                                            device=str(solver.device))
        self.ode_system = ode_system
        self.sampler = sampler
        self.n = int(n)
        self.solver = solver
        self.feature_extractor = feature_extractor
        # supervised (templates + 1-NN, or a classifier with fit/predict) or
        # unsupervised (a clusterer with fit_predict, e.g. HipHDBSCANClusterer)
        if template_y0 is None:
            if predictor is None or not hasattr(predictor, "fit_predict"):
                raise ValueError("without templates the predictor must be a clusterer (fit_predict)")
            self.template_y0 = torch.zeros((0, 0), dtype=torch.float32)
            self.template_labels = None
        else:
            self.template_y0 = torch.as_tensor(template_y0, dtype=torch.float32)
            self.template_labels = list(template_labels)
        self.predictor = predictor
        self.classifier = KNNTemplateClassifier(self.template_labels or [])
        self.keep_trajectories = keep_trajectories
        self.detect_unbounded = detect_unbounded
        self.compute_orbit_data = compute_orbit_data

    def run(self, y0: torch.Tensor | None = None) -> BSResult:
        return self._run(y0, None, [{"baseline": True}])[0]

    def run_parameter_study(self, study: list[dict], y0: torch.Tensor | None = None
                            ) -> list[BSResult]:
        """One result per study entry ({param name: value}), all B x P
        trajectories integrated in one launch, IC-major (row ic * P + p)."""
        grid = build_params_grid(self.ode_system, study)
        return self._run(y0, grid, list(study))

    # ------------------------------------------------------------------ #
    def _features(self, fx, res=None, solution=None) -> torch.Tensor:
        if res is not None and hasattr(fx, "extract_from_result"):
            return fx.extract_from_result(res)
        return fx.extract_features(solution)

    def _run(self, y0, grid, study_labels) -> list[BSResult]:
        if y0 is None:
            y0 = self.sampler.sample(self.n)
        B = y0.shape[0]
        P = 1 if grid is None else grid.shape[0]
        dev = self.solver.device
        main = y0.to(dev)
        if P > 1:
            main = torch.repeat_interleave(main, P, dim=0)
        tmpl = self.template_y0.to(dev, y0.dtype)
        if tmpl.numel() == 0:
            tmpl = tmpl.reshape(0, main.shape[1])
        batch = torch.cat([main, tmpl])
        params = None
        if grid is not None:
            base = build_params_grid(self.ode_system, [{}])
            params = torch.cat([grid.repeat(B, 1), base.expand(len(tmpl), -1)]).to(dev)
        fx = self.feature_extractor
        extra: dict = {}
        unbounded = None
        if dev.type == "cuda" and functor_id(self.ode_system) is not None:
            save_ts, batch_p = self.solver._prepare_tensors(batch)
            steady = getattr(fx, "time_steady", None)
            if self.compute_orbit_data:
                res, extra_orbit = self._orbit_rows(batch_p, save_ts, params, steady)
            else:
                res = self.solver.run_fused(self.ode_system, batch_p, save_ts, params,
                                            keep_trajectories=self.keep_trajectories,
                                            time_steady=steady,
                                            trend=bool(getattr(fx, "needs_trend", False)))
                extra_orbit = None
            if res.y is not None:
                FUSED_RESULTS[res.y] = res
            fused_ok = res.stats is not None and (
                isinstance(fx, HipFusedFeatureExtractor) or hasattr(fx, "extract_from_result"))
            if fused_ok:
                feats = self._features(fx, res=res)
            else:
                if res.y is None:
                    raise ValueError("a non-fused extractor needs keep_trajectories=True")
                feats = fx.extract_features(Solution(batch, save_ts, res.y))
            if self.detect_unbounded and self.solver.stop_threshold is not None \
                    and self.solver.inf_fill:
                unbounded = (res.status == _lib.BH_LANE_DIVERGED).cpu()
            extra = {"status": res.status, "nsteps": res.nsteps}
        else:
            t, y = self.solver.integrate(self.ode_system, batch, None) if params is None else \
                self._integrate_rows(batch, params)
            feats = fx.extract_features(Solution(batch, t, y))
            if self.detect_unbounded:
                unbounded = torch.isinf(y).any(dim=(0, 2)).cpu()
            extra_orbit = None
            if self.compute_orbit_data:
                from .orbit import extract_orbit_data

                extra_orbit = extract_orbit_data(t, y, self._orbit_dof(batch.shape[1]),
                                                 time_steady=getattr(fx, "time_steady", None))
        feats = feats.cpu() if feats.device.type != "cpu" else feats
        n_main = B * P
        supervised = self.template_labels is not None
        if supervised and self.predictor is not None:
            self.predictor.fit(feats[n_main:].numpy(), np.asarray(self.template_labels))
        elif supervised:
            self.classifier.fit(feats[n_main:])
        results = []
        for p, label in enumerate(study_labels):
            f_p = feats[:n_main][p::P]
            ub = unbounded[:n_main][p::P] if unbounded is not None else torch.zeros(B, dtype=torch.bool)
            labels = np.empty(B, dtype=object)
            labels[ub.numpy()] = "unbounded"
            if int(ub.sum()) < B:
                fb = f_p[~ub]
                if not supervised:
                    pred = self.predictor.fit_predict(fb.numpy())  # per group, as the reference
                elif self.predictor is not None:
                    pred = self.predictor.predict(fb.numpy())
                else:
                    pred = self.classifier.predict(fb.to(dev) if dev.type == "cuda" else fb)
                labels[(~ub).numpy()] = np.asarray(pred, dtype=object)
            bs = compute_bs(labels)
            names = list(getattr(fx, "feature_names", []))
            ex = dict(extra, group=p)
            if extra_orbit is not None:
                from .orbit import OrbitData

                ex["orbit_data"] = OrbitData(extra_orbit.peak_values[:, :n_main][:, p::P],
                                             extra_orbit.peak_counts[:n_main][p::P],
                                             extra_orbit.dof_indices, extra_orbit.time_steady)
            results.append(BSResult(bs, get_errors(bs, B), labels, f_p, names, y0, ex,
                                    study_label=label))
        return results

    def _orbit_dof(self, S: int) -> list[int]:
        c = self.compute_orbit_data
        return list(range(S)) if c is True else list(c)

    def _orbit_rows(self, batch_p, save_ts, params, steady):
        """Fused launch with peaks for already expanded per-row parameters."""
        from .orbit import orbit_data_from_result

        plan = self.solver.plan_fused(self.ode_system, batch_p, save_ts, params,
                                      time_steady=steady)
        cap = max(0, min(64, (len(save_ts) - plan.steady_idx - 1) // 2))
        res = self.solver.launch_fused(plan, keep_trajectories=self.keep_trajectories,
                                       peak_cap=cap)
        need = int(res.peak_counts.max()) if res.peak_counts.numel() else 0
        if need > cap:
            res = self.solver.launch_fused(plan, keep_trajectories=self.keep_trajectories,
                                           peak_cap=need)
        return res, orbit_data_from_result(res, self._orbit_dof(batch_p.shape[1]), steady)

    def _integrate_rows(self, batch, params):
        """CPU path with per-row parameters (rows already expanded)."""
        save_ts, b = self.solver._prepare_tensors(batch)
        return save_ts, self.solver._integrate(self.ode_system, b, save_ts, params)[1]
