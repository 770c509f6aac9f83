"""End-to-end labels of the HIP path against the reference's seeded ground truths.

The ICs are regenerated from seed 42 (bit-exact with the reference CSVs,
checked on the CPU in test_oracle_golden.py); the fixtures hold only the
labels (tests/integration/*/ground_truths, produced by the reference's own
pipeline with per-trajectory Dopri5 in float32).  Each case uses the
reference case study's feature and template recipe:

* pendulum case 1 (case_studies/pendulum/setup_pendulum_system.py:34-65):
  t in [0, 1000], 1000 saves, time_steady 950, log_delta of omega, 1-NN on
  templates [0.4, 0] FP / [2.7, 0] LC;
* Duffing (setup_duffing_oscillator_system.py:12-74): save window
  [900, 1000] with 5000 points, maximum and std of x, 1-NN on 5 templates;
* Lorenz (setup_lorenz_system.py:29-88): save window [900, 1000], 4000
  points, stop at |y|inf >= 200 with Inf fill (the Event semantics; such
  trajectories are "unbounded", basin_stability_estimator.py:519-527), mean x
  and 1-NN on the two bounded templates.
"""

import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pybasin_amd import _lib
from pybasin_amd.estimator import KNNTemplateClassifier, compute_bs
from pybasin_amd.features import feature_from_stats
from pybasin_amd.solvers import HipDopri5Solver
from pybasin_amd.systems import DuffingODE, LorenzODE, PendulumODE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(GOLD, "label_fixtures.json")) as f:
        return json.load(f)


def _report(name, got, want):
    got, want = np.asarray(got), np.asarray(want)
    agree = float((got == want).mean())
    bs_g, bs_w = compute_bs(got), compute_bs(want)
    n = len(want)
    worst = max(abs(bs_g.get(k, 0.0) - bs_w.get(k, 0.0)) / max(math.sqrt(bs_w[k] * (1 - bs_w[k]) / n), 1e-12)
                for k in bs_w)
    print(f"[result] {name}: label agreement {agree:.4f}; BS hip {bs_g} ref {bs_w}; "
          f"worst |dBS| / SE = {worst:.2f}", flush=True)
    return agree, worst


def _run(solver, ode, y0, templates, time_steady):
    batch = torch.cat([y0, torch.tensor(templates, dtype=y0.dtype)]).to("cuda")
    save_ts, batch = solver._prepare_tensors(batch)
    return solver.run_fused(ode, batch, save_ts, None, keep_trajectories=False, time_steady=time_steady)


def test_pendulum_case1_labels(cuda_device, fixtures):
    fx = fixtures["pendulum_case1"]
    y0 = torch.tensor(O.sample_uniform(fx["n"], fx["min_limits"], fx["max_limits"], fx["seed"]))
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cuda", cache_dir=None)
    res = _run(solver, PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}), y0, [[0.4, 0.0], [2.7, 0.0]], 950.0)
    feats = feature_from_stats("log_delta", res.stats)[:, 1:2]
    knn = KNNTemplateClassifier(["FP", "LC"]).fit(feats[-2:])
    lab = knn.predict(feats[:-2])
    agree, worst = _report("pendulum case1", lab, fx["labels"])
    assert agree >= 0.995
    assert worst <= 1.0  # every basin-stability fraction within its binomial SE


def test_duffing_labels(cuda_device, fixtures):
    fx = fixtures["duffing_main"]
    y0 = torch.tensor(O.sample_uniform(fx["n"], fx["min_limits"], fx["max_limits"], fx["seed"]))
    templates = [[-0.21, 0.02], [1.05, 0.77], [-0.67, 0.02], [-0.46, 0.30], [-0.43, 0.12]]
    names = ["$\\bar{y}_1$", "$\\bar{y}_2$", "$\\bar{y}_3$", "$\\bar{y}_4$", "$\\bar{y}_5$"]
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=5000, t_eval=(900.0, 1000.0), device="cuda",
                             cache_dir=None)
    res = _run(solver, DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), y0, templates, 900.0)
    feats = torch.stack([feature_from_stats("maximum", res.stats)[:, 0],
                         feature_from_stats("standard_deviation", res.stats)[:, 0]], 1)
    lab = KNNTemplateClassifier(names).fit(feats[-5:]).predict(feats[:-5])
    agree, worst = _report("duffing", lab, fx["labels"])
    assert agree >= 0.995
    assert worst <= 1.0


def test_lorenz_labels(cuda_device, fixtures):
    fx = fixtures["lorenz_main"]
    y0 = torch.tensor(O.sample_uniform(fx["n"], fx["min_limits"], fx["max_limits"], fx["seed"]))
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=4000, t_eval=(900.0, 1000.0), device="cuda",
                             cache_dir=None, stop_threshold=200.0, inf_fill=True)
    res = _run(solver, LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6}), y0,
               [[0.8, -3.0, 0.0], [-0.8, 3.0, 0.0]], 900.0)
    status = res.status.cpu().numpy()
    assert (status[-2:] == _lib.BH_LANE_OK).all()
    feats = feature_from_stats("mean", res.stats)[:, 0:1]
    lab = KNNTemplateClassifier(["chaotic attractor 1", "chaotic attractor 2"]).fit(feats[-2:]).predict(feats[:-2])
    lab = np.where(status[:-2] == _lib.BH_LANE_DIVERGED, "unbounded", lab)
    agree, worst = _report("lorenz", lab, fx["labels"])
    # the bounded attractors are chaotic, yet which one an IC reaches is
    # robust (boundary ICs are 0.1 % of this set: tests/test_hip_parity_full.py)
    assert ((lab == "unbounded") == (np.asarray(fx["labels"]) == "unbounded")).mean() >= 0.999
    assert agree >= 0.995
    assert worst <= 1.0


def _mcc(got, want):
    """Matthews correlation over the label set (multi-class, Gorodkin's form)."""
    labels = sorted(set(want) | set(got))
    idx = {l: i for i, l in enumerate(labels)}
    C = np.zeros((len(labels), len(labels)))
    for g, w in zip(got, want):
        C[idx[w], idx[g]] += 1
    t, p, c, s = C.sum(1), C.sum(0), np.trace(C), C.sum()
    den = math.sqrt((s * s - (p * p).sum()) * (s * s - (t * t).sum()))
    return 1.0 if den == 0 else float((c * s - (t * p).sum()) / den)


def test_pendulum_torque_sweep_parameter_study(cuda_device, fixtures):
    """Pendulum case 2: 20 torques x 2000 seeded ICs in one launch through
    run_parameter_study, against the reference's per-T label files
    (tests/integration/pendulum/ground_truths/case2); pass criterion of the
    reference's run_parameter_study_test (tests/integration/test_helpers.py:437-530):
    MCC >= 0.95 per parameter value."""
    from pybasin_amd.estimator import BasinStabilityEstimator
    from pybasin_amd.features import HipFusedFeatureExtractor

    fx = fixtures["pendulum_case2_T_sweep"]
    y0 = torch.tensor(O.sample_uniform(fx["n"], fx["min_limits"], fx["max_limits"], fx["seed"]))
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cuda", cache_dir=None)
    ext = HipFusedFeatureExtractor(time_steady=950.0, features=None,
                                   features_per_state={1: {"log_delta": None}})
    bse = BasinStabilityEstimator(PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}), None, fx["n"],
                                  solver, ext, template_y0=[[0.4, 0.0], [2.7, 0.0]],
                                  template_labels=["FP", "LC"])
    results = bse.run_parameter_study([{"T": e["T"]} for e in fx["sweep"]], y0)
    worst_mcc, worst_agree = 1.0, 1.0
    for r, e in zip(results, fx["sweep"]):
        agree = float((np.asarray(r.labels) == np.asarray(e["labels"])).mean())
        mcc = _mcc(list(r.labels), e["labels"])
        print(f"[result] T={e['T']:.2f} agreement {agree:.4f} MCC {mcc:.4f} BS {r.basin_stability}",
              flush=True)
        worst_mcc, worst_agree = min(worst_mcc, mcc), min(worst_agree, agree)
        assert mcc >= 0.95
    print(f"[result] torque sweep: worst agreement {worst_agree:.4f}, worst MCC {worst_mcc:.4f}")


def test_parameter_study_group_equals_single_launch(cuda_device):
    """Per-lane integration is independent of the batch: group p of a study
    equals a launch with that parameter set, bit for bit."""
    solver = HipDopri5Solver(t_span=(0, 50), t_steps=101, t_eval=(25.0, 50.0), device="cuda",
                             cache_dir=None)
    y0 = torch.tensor(O.sample_uniform(500, [-2.6, -10.0], [3.7, 10.0], 5)).to(cuda_device)
    grid = torch.tensor([[0.1, 0.3, 1.0], [0.1, 0.7, 1.0]], dtype=torch.float64)
    ode = PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0})
    both = solver.integrate_fused(ode, y0, grid, time_steady=30.0)
    for p in range(2):
        one = solver.integrate_fused(PendulumODE({"alpha": 0.1, "T": float(grid[p, 1]), "K": 1.0}),
                                     y0, time_steady=30.0)
        assert torch.equal(both.final[p::2], one.final)
        assert torch.equal(both.stats[p::2], one.stats)


def test_case_extractors_fused_equal_trajectory(cuda_device):
    """The case-study extractors give the same output from the kernel's fused
    statistics as from the kept trajectory of the same launch."""
    from pybasin_amd.case_features import (DuffingFeatureExtractor, LorenzFeatureExtractor,
                                           PendulumFeatureExtractor)
    from pybasin_amd.features import Solution
    from pybasin_amd.solvers.hip_dopri5 import FUSED_RESULTS

    cases = [
        (PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}), PendulumFeatureExtractor(time_steady=950.0),
         O.sample_uniform(400, [-2.6, -10.0], [3.7, 10.0], 1), {}),
        (DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), DuffingFeatureExtractor(time_steady=900.0),
         O.sample_uniform(400, [-1.0, -0.5], [1.0, 1.0], 2), {}),
        (LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6}), LorenzFeatureExtractor(time_steady=900.0),
         O.sample_uniform(400, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], 3), {"stop_threshold": 200.0}),
    ]
    for ode, ext, ics, kw in cases:
        solver = HipDopri5Solver(t_span=(0, 1000), t_steps=500, t_eval=(850.0, 1000.0),
                                 device="cuda", cache_dir=None, **kw)
        y0 = torch.tensor(ics).to(cuda_device)
        res = solver.integrate_fused(ode, y0, time_steady=ext.time_steady, keep_trajectories=True)
        fused = ext.extract_from_result(res)
        FUSED_RESULTS.pop(res.y, None)
        direct = ext.extract_features(Solution(y0, res.t, res.y))
        assert fused.shape == direct.shape
        torch.testing.assert_close(fused.cpu(), direct.cpu(), rtol=1e-5, atol=1e-6)


def test_estimator_orbit_data_per_group(cuda_device):
    """compute_orbit_data in a parameter study: group p's orbit data equals the
    trajectory peaks of that group's rows."""
    from pybasin_amd.estimator import BasinStabilityEstimator
    from pybasin_amd.features import HipFusedFeatureExtractor
    from pybasin_amd.orbit import extract_orbit_data

    y0 = torch.tensor(O.sample_uniform(300, [-2.6, -10.0], [3.7, 10.0], 4))
    solver = HipDopri5Solver(t_span=(0, 200), t_steps=400, t_eval=(100.0, 200.0), device="cuda",
                             cache_dir=None)
    ext = HipFusedFeatureExtractor(time_steady=150.0, features=None,
                                   features_per_state={1: {"log_delta": None}})
    ode = PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0})
    bse = BasinStabilityEstimator(ode, None, 300, solver, ext, template_y0=[[0.4, 0.0], [2.7, 0.0]],
                                  template_labels=["FP", "LC"], compute_orbit_data=[1])
    study = bse.run_parameter_study([{"T": 0.3}, {"T": 0.8}], y0)
    for r, T in zip(study, (0.3, 0.8)):
        t, y = solver.integrate(PendulumODE({"alpha": 0.1, "T": T, "K": 1.0}), y0.to("cuda"))
        want = extract_orbit_data(t, y, [1], time_steady=150.0)
        od = r.extra["orbit_data"]
        m = int(want.peak_counts.max())
        assert torch.equal(od.peak_counts, want.peak_counts)
        assert torch.equal(od.peak_values[:m].nan_to_num(-7.0), want.peak_values.nan_to_num(-7.0))


def test_estimator_with_the_reference_defaults(cuda_device):
    """BasinStabilityEstimator(ode, sampler, n) with nothing else takes the
    reference's defaults (basin_stability_estimator.py:73-181): Dopri5 on
    [0, 1000] saving 1,000 points of [850, 1000], the minimal feature set
    z-scored, and HDBSCAN with auto-tuned min_cluster_size and noise
    assignment; the labels equal those of the same pieces run by hand."""
    from pybasin_amd.cluster import HipHDBSCANClusterer
    from pybasin_amd.estimator import BasinStabilityEstimator, UniformRandomSampler

    T = 0.5
    sampler = UniformRandomSampler([-math.pi + math.asin(T), -10.0], [math.pi + math.asin(T), 10.0],
                                   device="cuda")
    torch.manual_seed(3)
    bse = BasinStabilityEstimator(PendulumODE({"alpha": 0.1, "T": T, "K": 1.0}), sampler, n=3000)
    assert bse.solver.t_span == (0, 1000) and tuple(bse.solver.t_eval) == (850.0, 1000.0)
    assert bse.keep_trajectories  # the median needs the saved window
    res = bse.run()
    assert len(res.feature_names) == 20 and res.feature_names[0] == "state_0__median"
    f = res.features.double()
    live = f.std(0) > 0
    assert torch.allclose(f.mean(0)[live], torch.zeros(int(live.sum()), dtype=torch.float64), atol=1e-5)
    assert abs(sum(res.basin_stability.values()) - 1.0) < 1e-9
    by_hand = HipHDBSCANClusterer(auto_tune=True, assign_noise=True).fit_predict(res.features.numpy())
    assert np.array_equal(np.asarray(res.labels, dtype=np.int64), by_hand)
