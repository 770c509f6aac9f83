"""CPU tests of the host side: the torch shared-step Dopri5 (fallback and CPU
clone), the Solver contract on it, feature extraction and the BS driver."""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pybasin_amd.estimator import (BasinStabilityEstimator, KNNTemplateClassifier,
                                   UniformRandomSampler, compute_bs, get_errors)
from pybasin_amd.features import (HipFusedFeatureExtractor, Solution, StatFeatureExtractor,
                                  feature_from_stats, impute_extreme, stats_from_window)
from pybasin_amd.solvers import HipDopri5Solver, TorchDopri5Solver, functor_id, register_functor
from pybasin_amd.systems import DuffingODE, ExponentialDecayODE, LorenzODE, PendulumODE
from pybasin_amd.ode_system import ODESystem


@pytest.mark.parametrize("dtype,bits", [(torch.float32, 32), (torch.float64, 64)])
def test_torch_solver_equals_oracle_shared_mode(dtype, bits):
    """The torch fallback runs torchdiffeq's shared-dt algorithm: it must track
    the oracle's shared mode on a small batch (same decisions, rounding only)."""
    y0 = torch.tensor(O.sample_uniform(6, [-1.0, -0.5], [1.0, 1.0], 9), dtype=dtype)
    s = TorchDopri5Solver(t_span=(0, 20), t_steps=41, device="cpu", cache_dir=None)
    t, y = s.integrate(DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), y0)
    ref, st, _ = O.solve(O.SYS_DUFFING, y0.double().numpy(), np.array([0.08, 1.0, 0.2]),
                         t.double().numpy(), 0.0, bits=bits, mode=O.MODE_TD_SHARED)
    tol = 1e-4 if bits == 32 else 1e-9
    np.testing.assert_allclose(y.double().numpy(), ref, atol=tol, rtol=0)


def test_torch_solver_reproduces_reference_torchdiffeq_table():
    """Shared-step float64 run of the 5 zigode test ICs: final theta equals the
    reference's published torchdiffeq column (src/zigode/README.md:70-76)."""
    ics = torch.tensor([[0.4, 0.0], [2.7, 0.0], [0.5, -5.0], [-2.0, 3.0], [1.0, 0.0]], dtype=torch.float64)
    s = TorchDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cpu", cache_dir=None)
    _, y = s.integrate(PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}), ics)
    want = [0.523599, 4947.703034, 4904.614031, 4975.223032, 0.523599]
    np.testing.assert_allclose(y[-1, :, 0].numpy(), want, atol=2e-5)


def test_solver_contract_on_cpu():
    ode = ExponentialDecayODE({"decay": -1.0})
    s = TorchDopri5Solver(t_span=(0, 1), t_steps=10, device="cpu", cache_dir=None)
    t, y = s.integrate(ode, torch.tensor([[1.0]]))
    assert t.shape == (10,) and y.shape == (10, 1, 1)
    assert y[0].item() == pytest.approx(1.0, abs=1e-5)
    assert y[-1].item() == pytest.approx(math.exp(-1), abs=1e-5)
    s2 = TorchDopri5Solver(t_span=(0, 2), t_steps=10, t_eval=(1, 2), device="cpu", cache_dir=None)
    _, y2 = s2.integrate(ode, torch.tensor([[1.0]]))
    assert y2[0].item() == pytest.approx(math.exp(-1), abs=1e-5)
    _, y3 = TorchDopri5Solver(t_span=(0, 1), t_steps=11, device="cpu", cache_dir=None).integrate(
        ode, torch.tensor([[1.0]]), params=torch.tensor([[-1.0], [-2.0]]))
    assert y3.shape == (11, 2, 1)
    assert y3[-1, 1, 0].item() == pytest.approx(math.exp(-2), abs=1e-5)
    with pytest.raises(ValueError):
        s.integrate(ode, torch.tensor([1.0]))
    with pytest.raises(ValueError, match="t_eval"):
        TorchDopri5Solver(t_span=(0, 1), t_eval=(0.5, 1.5), device="cpu", cache_dir=None)


def test_hip_solver_cpu_clone_uses_torch_path():
    """A HIP solver on the CPU (TemplateIntegrator's clone) integrates with the
    torch shared-step algorithm, no GPU needed."""
    s = HipDopri5Solver(t_span=(0, 1), t_steps=5, device="cpu", cache_dir=None)
    t, y = s.integrate(ExponentialDecayODE({"decay": -1.0}), torch.tensor([[1.0]]))
    assert y[-1].item() == pytest.approx(math.exp(-1), abs=1e-5)
    c = HipDopri5Solver(t_span=(0, 2), t_eval=(1, 2), device="cuda", cache_dir=None).clone(device="cpu")
    assert c.device.type == "cpu" and c.t_eval == (1, 2)


def test_functor_registry_is_explicit():
    class MyPendulum(ODESystem):
        def ode(self, t, y, p):
            return torch.zeros_like(y)

    assert functor_id(MyPendulum({"a": 1.0})) is None
    register_functor(MyPendulum, 2)
    assert functor_id(MyPendulum({"a": 1.0})) == 2
    assert functor_id(PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0})) == 2


def test_cache_roundtrip(tmp_path):
    s = TorchDopri5Solver(t_span=(0, 1), t_steps=5, device="cpu", cache_dir=str(tmp_path))
    ode = ExponentialDecayODE({"decay": -1.0})
    t1, y1 = s.integrate(ode, torch.tensor([[1.0]]))
    files = list(tmp_path.iterdir())
    assert len(files) == 1
    t2, y2 = s.integrate(ode, torch.tensor([[1.0]]))
    assert torch.equal(y1, y2)
    files[0].write_bytes(b"corrupt")
    t3, y3 = s.integrate(ode, torch.tensor([[1.0]]))
    assert torch.allclose(y3, y1)


def test_feature_columns_and_imputation():
    t = torch.linspace(0, 10, 101)
    y = torch.stack([torch.sin(t), t, torch.full_like(t, 2.0)], -1)[:, None, :].repeat(1, 3, 1)
    sol = Solution(y[0], t, y)
    fx = StatFeatureExtractor(time_steady=5.0, features=None,
                              features_per_state={0: {"maximum": None, "standard_deviation": None},
                                                  2: {"log_delta": None}})
    f = fx.extract_features(sol)
    assert fx.feature_names == ["state_0__maximum", "state_0__standard_deviation", "state_2__log_delta"]
    w = O.features(y[50:].double().numpy())
    np.testing.assert_allclose(f[:, 0].numpy(), w["maximum"][:, 0], rtol=1e-6)
    np.testing.assert_allclose(f[:, 1].numpy(), w["standard_deviation"][:, 0], rtol=1e-5)
    assert f[0, 2].item() == pytest.approx(math.log(1e-10), rel=1e-5)
    # the fused extractor falls back to the trajectory when no kernel stats exist
    f2 = HipFusedFeatureExtractor(time_steady=5.0, features=None,
                                  features_per_state={0: {"maximum": None, "standard_deviation": None},
                                                      2: {"log_delta": None}}).extract_features(sol)
    assert torch.equal(f, f2)
    imp = impute_extreme(torch.tensor([[float("inf"), -float("inf"), float("nan"), 3.0]]))
    assert imp.tolist() == [[1e10, -1e10, 1e10, 3.0]]
    st = stats_from_window(y[50:])
    assert torch.allclose(feature_from_stats("delta", st), (st[..., 3] - st[..., 0]).abs())


def test_bs_and_errors_match_reference_json():
    labels = np.array(["FP"] * 602 + ["LC"] * 3398)
    bs = compute_bs(labels)
    err = get_errors(bs, len(labels))
    assert bs["FP"] == pytest.approx(0.1505)
    assert err["FP"]["e_abs"] == pytest.approx(0.005653533187308623)


def test_estimator_on_cpu_with_templates():
    """integrate -> features -> 1-NN on templates -> BS, on the CPU path."""
    torch.manual_seed(0)
    sampler = UniformRandomSampler([-1.0, -0.5], [1.0, 1.0], device="cpu")
    solver = TorchDopri5Solver(t_span=(0, 20), t_steps=21, device="cpu", cache_dir=None, rtol=1e-6)
    fx = StatFeatureExtractor(time_steady=10.0, features=None,
                              features_per_state={0: {"maximum": None}})
    bse = BasinStabilityEstimator(ExponentialDecayODE({"decay": -1.0}), sampler, 8, solver, fx,
                                  template_y0=[[0.5, 0.5], [-0.5, -0.5]], template_labels=["pos", "neg"])
    y0 = torch.tensor([[0.9, 0.1], [-0.9, 0.1], [0.4, -0.4], [-0.2, 0.3]])
    res = bse.run(y0)
    assert list(res.labels) == ["pos", "neg", "pos", "neg"]
    assert res.basin_stability == {"neg": 0.5, "pos": 0.5}


def test_knn_template_classifier():
    knn = KNNTemplateClassifier(["a", "b"]).fit(torch.tensor([[0.0], [1.0]]))
    assert list(knn.predict(torch.tensor([[0.2], [0.9], [-5.0]]))) == ["a", "b", "a"]


def test_case_study_extractors_follow_reference_columns():
    """Pendulum / Duffing / Lorenz extractors (case_studies/*_feature_extractor.py)
    on synthetic windows, against direct torch reductions of the window."""
    from pybasin_amd.case_features import (DuffingFeatureExtractor, LorenzFeatureExtractor,
                                           PendulumFeatureExtractor)

    t = torch.linspace(0, 10, 101)
    # pendulum: omega constant (FP) / oscillating (LC)
    om_fp = torch.full_like(t, 0.3)
    om_lc = 1.0 + torch.sin(3 * t)
    y = torch.stack([torch.stack([t, om_fp], -1), torch.stack([t, om_lc], -1)], 1)
    sol = Solution(y[0], t, y)
    f = PendulumFeatureExtractor(time_steady=5.0).extract_features(sol)
    assert f.dtype == torch.float64 and f.tolist() == [[1.0, 0.0], [0.0, 1.0]]
    # Duffing: max and ddof-1 std of x over the window
    x = torch.stack([torch.sin(t) * 2, torch.cos(0.5 * t)], 1)
    yd = torch.stack([x, torch.zeros_like(x)], -1)
    fd = DuffingFeatureExtractor(time_steady=5.0).extract_features(Solution(yd[0], t, yd))
    w = yd[50:, :, 0]
    assert torch.allclose(fd[:, 0], w.max(0).values) and fd.dtype == torch.float32
    assert torch.allclose(fd[:, 1], torch.std(w, dim=0), rtol=1e-5)
    # Lorenz: sign of mean x, [0, 0] once any state exceeds 195 in the window
    xl = torch.stack([torch.full_like(t, 2.0), torch.full_like(t, -2.0), torch.full_like(t, 2.0)], 1)
    zl = torch.zeros_like(xl)
    zl[80, 2] = -196.0
    yl = torch.stack([xl, zl, zl], -1)
    fl = LorenzFeatureExtractor(time_steady=5.0).extract_features(Solution(yl[0], t, yl))
    assert fl.tolist() == [[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]]


def test_parameter_study_groups_match_single_runs_on_cpu():
    """run_parameter_study integrates B x P IC-major rows in one call; every
    group equals a separate run with that parameter set."""
    from pybasin_amd.estimator import build_params_grid

    ode = ExponentialDecayODE({"decay": -1.0})
    g = build_params_grid(ode, [{"decay": -0.5}, {}, {"decay": 0.5}])
    assert g.tolist() == [[-0.5], [-1.0], [0.5]]
    solver = TorchDopri5Solver(t_span=(0, 4), t_steps=9, device="cpu", cache_dir=None)
    fx = StatFeatureExtractor(time_steady=2.0, features=None,
                              features_per_state={0: {"mean": None}})
    y0 = torch.tensor([[0.9], [-0.9], [0.4], [-0.2]])
    bse = BasinStabilityEstimator(ode, None, 4, solver, fx,
                                  template_y0=[[0.5], [-0.5]], template_labels=["pos", "neg"])
    study = bse.run_parameter_study([{"decay": -0.5}, {"decay": 0.5}], y0)
    assert [r.study_label for r in study] == [{"decay": -0.5}, {"decay": 0.5}]
    for r, d in zip(study, (-0.5, 0.5)):
        single = BasinStabilityEstimator(ExponentialDecayODE({"decay": d}), None, 4, solver, fx,
                                         template_y0=[[0.5], [-0.5]],
                                         template_labels=["pos", "neg"])
        # templates use the base parameters in the study, so compare features only
        assert torch.allclose(r.features, single.run(y0).features, rtol=1e-6)
        assert list(r.labels) == ["pos", "neg", "pos", "neg"]


def test_unbounded_rows_are_labelled_and_excluded_on_cpu():
    """Inf in a trajectory marks the row unbounded (basin_stability_estimator.py:519-527);
    unbounded rows are left out of classification and counted in BS."""
    class InfAfterEscape(TorchDopri5Solver):
        def _integrate(self, ode_system, y0, save_ts, params=None):
            t, y = super()._integrate(ode_system, y0, save_ts, params)
            y = y.clone()
            y[2:, y0[:, 0].abs() > 5] = float("inf")  # Event semantics: Inf after the stop
            return t, y

    solver = InfAfterEscape(t_span=(0, 1), t_steps=5, device="cpu", cache_dir=None)
    fx = StatFeatureExtractor(time_steady=0.5, features=None, features_per_state={0: {"mean": None}})
    bse = BasinStabilityEstimator(ExponentialDecayODE({"decay": -1.0}), None, 4, solver, fx,
                                  template_y0=[[0.5], [-0.5]], template_labels=["pos", "neg"])
    res = bse.run(torch.tensor([[0.9], [9.0], [-0.4], [-9.0]]))
    assert list(res.labels) == ["pos", "unbounded", "neg", "unbounded"]
    assert res.basin_stability == {"neg": 0.25, "pos": 0.25, "unbounded": 0.5}


def test_peak_mask_matches_scipy_argrelmax():
    """find_peak_mask is strict local maxima with support n (scipy.signal.argrelmax,
    which the reference names as its equivalent, torch_features_pattern.py:64-78)
    on the interior; plateaus give no peak and the first and last n points are
    excluded (argrelmax's default clip mode can report those)."""
    from scipy.signal import argrelmax

    from pybasin_amd.orbit import extract_orbit_data, extract_peak_values, find_peak_mask

    g = torch.Generator().manual_seed(0)
    x = torch.randint(0, 6, (64, 5, 3), generator=g).float()  # many ties
    for n in (1, 2, 3):
        m = find_peak_mask(x, n)
        for b in range(5):
            for s in range(3):
                want = np.zeros(64, bool)
                want[argrelmax(x[:, b, s].numpy(), order=n)[0]] = True
                want[:n] = False
                want[64 - n:] = False
                assert (m[:, b, s].numpy() == want).all()
    vals, counts = extract_peak_values(x, 1)
    for b in range(5):
        for s in range(3):
            pk = x[1:-1, b, s][find_peak_mask(x, 1)[1:-1, b, s]]
            assert counts[b, s] == len(pk)
            assert torch.equal(vals[: len(pk), b, s], pk)
            assert torch.isnan(vals[len(pk):, b, s]).all()
    t = torch.linspace(0, 20, 201)
    y = torch.stack([torch.sin(t), torch.cos(2 * t)], -1)[:, None, :]
    od = extract_orbit_data(t, y, dof=[1], time_steady=10.0)
    assert od.peak_counts.tolist() == [[3]] and od.time_steady == 10.0
    # samples every 0.1: a sampled maximum of cos(2t) is within 1 - cos(0.1) of 1
    assert torch.allclose(od.peak_values[:, 0, 0], torch.ones(3), atol=1 - math.cos(0.1))


def test_window_features_follow_torch_calculators_on_cpu():
    """median = x.median(dim=0).values (lower middle), quantile =
    torch.quantile, root_mean_square = sqrt(mean(x^2)) (torch_features_statistical.py),
    names state_{i}__quantile__q_0.1 (feature_extractors/utils.py:34-80)."""
    from pybasin_amd.features import MINIMAL_FEATURES, feature_name, window_feature

    g = torch.Generator().manual_seed(1)
    for N in (7, 8):
        y = torch.randint(-3, 4, (N, 5, 2), generator=g).to(torch.float32)  # ties
        y[0, 0, 0] = float("nan")
        got = window_feature("median", None, y, 2)
        want = y[2:].median(dim=0).values
        assert torch.equal(got.nan_to_num(99), want.nan_to_num(99))
        for q in (0.1, 0.5, 0.9):
            got = window_feature("quantile", {"q": q}, y, 1)
            want = torch.quantile(y[1:], q, dim=0)
            assert torch.allclose(got.nan_to_num(99), want.nan_to_num(99), atol=1e-6)
    assert feature_name(1, "quantile", {"q": 0.1}) == "state_1__quantile__q_0.1"
    t = torch.linspace(0, 10, 101)
    yy = torch.stack([torch.sin(t), t, torch.cos(3 * t)], -1)[:, None, :].repeat(1, 2, 1)
    fx = StatFeatureExtractor(time_steady=5.0, features=MINIMAL_FEATURES)
    f = fx.extract_features(Solution(yy[0], t, yy))
    w = yy[50:].double()
    ref = {"median": w.median(dim=0).values, "mean": w.mean(0), "standard_deviation": w.std(0, correction=0),
           "variance": w.var(0, correction=0), "root_mean_square": torch.sqrt((w ** 2).mean(0)),
           "maximum": w.max(0).values, "absolute_maximum": w.abs().max(0).values,
           "minimum": w.min(0).values, "delta": (w.max(0).values - w.mean(0)).abs(),
           "log_delta": torch.log((w.max(0).values - w.mean(0)).abs() + 1e-10)}
    assert fx.feature_names[:10] == [f"state_0__{k}" for k in MINIMAL_FEATURES]
    for j, name in enumerate(fx.feature_names):
        s, feat = int(name.split("__")[0][6:]), name.split("__")[1]
        assert f[0, j].item() == pytest.approx(ref[feat][0, s].item(), rel=1e-5, abs=1e-6), name


def test_duffing_stage_sincos_coefficients_are_within_bound():
    """The float polynomial that the Duffing functor uses for sin/cos of a
    stage offset |d| <= 1 (coefficients read from systems.hpp), evaluated
    in float32 in the kernel's order, stays within 1.3 ulp of sin/cos."""
    import os
    import re

    src = open(os.path.join(os.path.dirname(__file__), "..", "pybasin_amd", "csrc",
                            "systems.hpp")).read()
    body = src[src.index("sincos_unit<float>(float d"):]
    body = body[:body.index("}")]
    lit = [np.float32(v) for v in re.findall(r"(-?\d\.\d+e[-+]\d+)f", body)]
    assert len(lit) == 7, lit
    p, q = lit[:3], lit[3:]
    f = np.float32
    d = np.linspace(-1, 1, 400001).astype(f)
    x = (d * d).astype(f)

    def horner(cs):
        r = cs[-1]
        for c in cs[-2::-1]:
            r = (c + x * r).astype(f)
        return r

    s = (d + (d * x).astype(f) * horner(p)).astype(f)
    c = (f(1) - x * horner(q)).astype(f)
    dd = d.astype(np.float64)
    err_s = np.abs(s - np.sin(dd)) / np.spacing(np.abs(np.sin(dd)).astype(f))
    err_c = np.abs(c - np.cos(dd)) / np.spacing(np.cos(dd).astype(f))
    assert err_s.max() <= 1.3 and err_c.max() <= 1.3, (err_s.max(), err_c.max())
    # a dropped or sign-flipped term is far outside the bound
    bad = (d + (d * x).astype(f) * horner([p[0], p[1], -p[2]])).astype(f)
    assert (np.abs(bad - np.sin(dd)) / np.spacing(np.abs(np.sin(dd)).astype(f))).max() > 10


def test_fused_result_lookup_by_tensor_identity():
    """A fused result registered for a trajectory tensor is found again from a
    Solution holding that same tensor (the reference estimator passes y on
    unchanged, basin_stability_estimator.py:282,307), and a tensor with equal
    values is a different key."""
    from pybasin_amd.case_features import PendulumFeatureExtractor
    from pybasin_amd.solvers import FUSED_RESULTS, FusedResult

    t = torch.linspace(0, 10, 101)
    y = torch.stack([torch.sin(t), torch.cos(t)], -1)[:, None, :].repeat(1, 4, 1).contiguous()
    stats = stats_from_window(y[50:]).clone()
    stats[:, 0, 3] = 7.0  # a marker the trajectory itself does not give
    res = FusedResult(t=t, y=y, final=y[-1], stats=stats, status=torch.zeros(4, dtype=torch.int32),
                      nsteps=torch.ones(4, dtype=torch.int64), steady_idx=50)
    FUSED_RESULTS[y] = res
    other = y.clone()
    fx = HipFusedFeatureExtractor(time_steady=5.0, features=None,
                                  features_per_state={0: {"maximum": None}})
    f = fx.extract_features(Solution(y[0], t, y))
    assert torch.equal(f[:, 0], torch.full((4,), 7.0))
    f_other = fx.extract_features(Solution(other[0], t, other))
    assert torch.allclose(f_other[:, 0], y[50:, :, 0].max(0).values)
    assert FUSED_RESULTS.get(other) is None
    # the case-study extractors look their results up the same way
    PendulumFeatureExtractor(time_steady=5.0).extract_features(Solution(y[0], t, y))
    del FUSED_RESULTS[y]


class _NoParamStrODE(ODESystem):
    """An ODE whose description carries no parameter values (as the
    reference's LorenzODE.get_str, lorenz_ode.py:66-72)."""

    def ode(self, t, y, p):
        return p[..., 0:1] * y

    def get_str(self):
        return "decay without parameters in its description"


def test_cache_key_covers_ode_parameters(tmp_path):
    s = TorchDopri5Solver(t_span=(0, 1), t_steps=5, device="cpu", cache_dir=str(tmp_path))
    y0 = torch.tensor([[1.0]])
    _, ya = s.integrate(_NoParamStrODE({"decay": -1.0}), y0)
    _, yb = s.integrate(_NoParamStrODE({"decay": -2.0}), y0)
    assert len(list(tmp_path.iterdir())) == 2
    assert not torch.allclose(ya, yb)
    _, ya2 = s.integrate(_NoParamStrODE({"decay": -1.0}), y0)
    assert torch.equal(ya, ya2) and len(list(tmp_path.iterdir())) == 2


def test_zscore_keeps_constant_columns_at_unit_scale():
    """A column constant in the first batch keeps std 1 (the reference's
    torch.where(std == 0, 1, std), torch_feature_extractor.py:273-281), so a
    later batch that differs there is not scaled up."""
    t = torch.linspace(0, 10, 101)
    y = torch.stack([torch.sin(t), torch.full_like(t, 2.0)], -1)[:, None, :].repeat(1, 3, 1)
    y[:, 1, 0] *= 2
    fx = StatFeatureExtractor(time_steady=5.0, features={"mean": None}, normalize=True)
    fx.extract_features(Solution(y[0], t, y))
    y2 = y.clone()
    y2[:, :, 1] = 3.0
    f2 = fx.extract_features(Solution(y2[0], t, y2))
    assert torch.allclose(f2[:, 1], torch.full((3,), 1.0))


def test_dynamical_system_feature_set_on_cpu_equals_oracle():
    """The reference's dynamical-system set (settings.py:407-417) through the
    CPU path of StatFeatureExtractor, column by column against the oracle's
    restatements (each pinned by closed forms in test_oracle_golden.py)."""
    from pybasin_amd.features import DYNAMICAL_SYSTEM_FEATURES

    g = torch.Generator().manual_seed(0)
    t = torch.linspace(0, 60, 600)
    B = 6
    amp = torch.rand(B, generator=g) + 0.5
    per = torch.tensor([3.0, 5.0, 7.0, 2.0, 11.0, 4.0])
    y0 = torch.stack([amp[None, :] * torch.sin(2 * math.pi * t[:, None] / per[None, :]) + 0.01 * t[:, None],
                      torch.cos(2 * math.pi * t[:, None] / 6.0).expand(600, B) * amp[None, :]], -1)
    y = y0 + 0.01 * torch.randn(600, B, 2, generator=g)
    sol = Solution(y[0], t, y)
    fx = StatFeatureExtractor(time_steady=20.0, features=DYNAMICAL_SYSTEM_FEATURES)
    f = fx.extract_features(sol).double()
    assert fx.feature_names[:7] == [
        "state_0__variance", "state_0__amplitude", "state_0__mean", "state_0__linear_trend__attr_slope",
        "state_0__autocorrelation_periodicity__output_strength",
        "state_0__autocorrelation_periodicity__output_period", "state_0__spectral_frequency_ratio"]
    start = int(torch.nonzero(t >= 20.0)[0])
    w = y[start:].double().numpy()
    strength, period = O.autocorrelation_periodicity(w)
    want = {"variance": w.var(0), "amplitude": O.amplitude(w), "mean": w.mean(0),
            "linear_trend__attr_slope": O.linear_trend_slope(w),
            "autocorrelation_periodicity__output_strength": strength,
            "autocorrelation_periodicity__output_period": period,
            "spectral_frequency_ratio": O.spectral_frequency_ratio(w)}
    for j, name in enumerate(fx.feature_names):
        s, feat = int(name.split("__")[0][6:]), name.split("__", 1)[1]
        np.testing.assert_allclose(f[:, j].numpy(), want[feat][:, s].astype(np.float32),
                                   rtol=2e-5, atol=2e-6, err_msg=name)
    assert fx.needs_window and fx.needs_trend


def test_float_sine_cosine_are_within_bound():
    """The branch-free float sine and cosine of systems.hpp (Cody-Waite
    reduction by pi/2 in three FMA steps, polynomials on |r| <= 0.8,
    quadrant selection; constants read from the source), evaluated in
    float32 in the kernel's order with exact FMAs, stay within 7.5e-8 of
    sin x and cos x for |x| <= 2^17, near multiples of pi/2 included; a
    dropped reduction step or polynomial term, or a quadrant sign taken from
    the wrong bit, breaks the bound."""
    import os
    import re

    src = open(os.path.join(os.path.dirname(__file__), "..", "pybasin_amd", "csrc",
                            "systems.hpp")).read()

    def body(sig):
        b = src[src.index(sig):]
        return b[:b.index("\n}")]

    lit = []
    for sig in ("void cw_reduce(", "float cw_sin_poly(", "float cw_cos_poly("):
        lit += [np.float32(v) for v in re.findall(r"(-?\d\.\d+e[-+]\d+)f", body(sig))]
    assert len(lit) == 11, lit
    two_pi_inv, c1, c2, c3, s2, s1, s0, k3, k2, k1, k0 = lit
    assert re.search(r"a = \(q & 1\) \? cv : sv, b = \(q & 1\) \? sv : cv", body("void sincos_fast("))
    assert "c = ((q + 1) & 2) ? -b : b" in body("void sincos_fast(")
    f = np.float32

    def fma(a, b, c):
        return (np.float64(a) * np.float64(b) + np.float64(c)).astype(f)

    def sincos(x, drop=None, cos_bit=1):
        k = np.rint((x * two_pi_inv).astype(f)).astype(f)
        r = fma(-k, c1, x)
        r = fma(-k, c2, r) if drop != "c2" else r
        r = fma(-k, c3, r)
        r2 = (r * r).astype(f)
        ps = fma(r2, fma(r2, s2 if drop != "s2" else f(0), s1), s0)
        sv = fma((r2 * r).astype(f), ps, r)
        cv = fma(r2, fma(r2, fma(r2, fma(r2, k3, k2), k1), k0), f(1))
        q = k.astype(np.int64)
        a, b = np.where(q & 1, cv, sv), np.where(q & 1, sv, cv)
        return np.where(q & 2, -a, a), np.where((q + cos_bit) & 2, -b, b)

    assert k0 == f(-0.5)
    rng = np.random.default_rng(3)
    x = rng.uniform(-131072, 131072, 400_000).astype(f)
    m = (rng.integers(-80_000, 80_000, 100_000) * np.pi / 2).astype(f)
    x = np.concatenate([x, m, np.nextafter(m, f(np.inf)), rng.uniform(-4, 4, 100_000).astype(f)])
    ts, tc = np.sin(x.astype(np.float64)), np.cos(x.astype(np.float64))
    s, c = sincos(x)
    assert np.abs(s - ts).max() <= 7.5e-8
    assert np.abs(c - tc).max() <= 7.5e-8
    for d in ("c2", "s2"):
        assert np.abs(sincos(x, drop=d)[0] - ts).max() > 1e-6, d
    assert np.abs(sincos(x, cos_bit=0)[1] - tc).max() > 0.5


def test_float_sine_by_pi_reduction_is_within_bound():
    """The pendulum's branch-free float sine (systems.hpp sin_fast:
    Cody-Waite reduction by pi in three FMA steps, the parity of k read from
    the last bit of x / pi + 1.5 * 2^23, the sign applied to r, an odd
    polynomial of degree 9; constants read from the source), evaluated in
    float32 in the kernel's order with exact FMAs, stays within 1.3e-7 of
    sin x for |x| <= 2^17, near multiples of pi/2 and pi included; a dropped
    reduction step or polynomial term, or the sign taken from the wrong bit,
    breaks the bound."""
    import os
    import re

    src = open(os.path.join(os.path.dirname(__file__), "..", "pybasin_amd", "csrc",
                            "systems.hpp")).read()
    b = src[src.index("float sin_fast(float x) {"):]
    b = b[:b.index("\n}")]
    lit = [np.float32(v) for v in re.findall(r"(-?\d\.\d+e[-+]\d+)f", b)]
    assert len(lit) == 10, lit
    inv_pi, magic, magic2, p1, p2, p3, c9, c7, c5, c3 = lit
    assert magic == magic2 == np.float32(1.5 * 2**23)
    assert "(__float_as_int(t) << 31)" in b
    f = np.float32

    def fma(a, b, c):
        return (np.float64(a) * np.float64(b) + np.float64(c)).astype(f)

    def sin32(x, drop=None, parity_bit=0):
        t = fma(x, inv_pi, magic)
        k = (t - magic).astype(f)
        r = fma(-k, p1, x)
        r = fma(-k, p2, r) if drop != "p2" else r
        r = fma(-k, p3, r)
        par = (t.view(np.int32) >> parity_bit) & 1
        r = (r.view(np.int32) ^ (par << 31)).view(f)
        r2 = (r * r).astype(f)
        p = fma(fma(fma(c9 if drop != "c9" else f(0), r2, c7), r2, c5), r2, c3)
        return fma((r2 * r).astype(f), p, r)

    rng = np.random.default_rng(5)
    x = rng.uniform(-131072, 131072, 400_000).astype(f)
    m = rng.integers(-80_000, 80_000, 100_000)
    near = np.concatenate([(m * np.pi / 2).astype(f), (m * np.pi / 2 / 2).astype(f)])
    x = np.concatenate([x, near, np.nextafter(near, f(np.inf)), np.nextafter(near, f(-np.inf)),
                        rng.uniform(-4, 4, 100_000).astype(f)])
    x = x[np.abs(x) <= 131072]
    ts = np.sin(x.astype(np.float64))
    assert np.abs(sin32(x) - ts).max() <= 1.3e-7
    for d in ("p2", "c9"):
        assert np.abs(sin32(x, drop=d) - ts).max() > 1e-6, d
    assert np.abs(sin32(x, parity_bit=1) - ts).max() > 0.5


def test_rossler_edge_arrays_give_the_laplacian():
    """The edge arrays handed to bh_solve_rossler_edges (both directions of
    every edge, the reference's build_edge_arrays) reproduce L = D - A of the
    dense kernel: sum over e with edges_i[e] = i of (x_i - x_edges_j[e]) equals
    (L x)_i for random x, and no node has more than the kernel's 12 entries."""
    from pybasin_amd.systems.rossler_network import _laplacian_np, rossler_edge_arrays

    ei, ej = rossler_edge_arrays()
    assert ei.dtype == np.int32 and len(ei) == 800 and ei.flags.c_contiguous
    x = np.random.default_rng(0).normal(size=(5, 100))
    cpl = np.zeros_like(x)
    np.add.at(cpl.T, ei, (x[:, ei] - x[:, ej]).T)
    np.testing.assert_allclose(cpl, x @ _laplacian_np().T.astype(np.float64), atol=1e-12)
    assert np.bincount(ei, minlength=100).max() <= 12
