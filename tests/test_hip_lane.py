"""GPU parity of the per-lane HIP Dopri5 kernel (libbasinhip.so) against the CPU oracle.

The oracle's per-lane torchdiffeq mode applies the same step rules to each
IC; the kernel differs only in rounding (FMA contraction, device sin/cos,
single-precision controller power), so trajectories agree to a small multiple
of the tolerance on short horizons, and the fused statistics must equal the
statistics of the kernel's own written trajectory.
"""

import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pybasin_amd import _lib
from pybasin_amd.solvers import HipDopri5Solver
from pybasin_amd.systems import DuffingODE, ExponentialDecayODE, Linear2ODE, LorenzODE, PendulumODE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")

PEND = {"alpha": 0.1, "T": 0.5, "K": 1.0}
DUFF = {"delta": 0.08, "k3": 1.0, "A": 0.2}
LOR = {"sigma": 0.12, "r": 0.0, "b": -0.6}

CASES = {
    # name: (ode, oracle sys, lo, hi, t_end)
    "pendulum": (PendulumODE(PEND), O.SYS_PENDULUM, [-math.pi + math.asin(0.5), -10.0],
                 [math.pi + math.asin(0.5), 10.0], 20.0),
    "duffing": (DuffingODE(DUFF), O.SYS_DUFFING, [-1.0, -0.5], [1.0, 1.0], 20.0),
    "lorenz": (LorenzODE(LOR), O.SYS_LORENZ, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], 10.0),
    "linear2": (Linear2ODE({"m00": -0.1, "m01": 1.0, "m10": -1.0, "m11": -0.1}), O.SYS_LINEAR2,
                [-1.0, -1.0], [1.0, 1.0], 20.0),
}


def _oracle_params(ode):
    return np.array([float(v) for v in ode.params.values()])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("name", list(CASES))
def test_trajectories_match_oracle_short_horizon(cuda_device, name, dtype):
    ode, sysid, lo, hi, t_end = CASES[name]
    print(f"[case] {name} {dtype}", flush=True)
    y0 = torch.tensor(O.sample_uniform(64, lo, hi, seed=7), dtype=dtype)
    solver = HipDopri5Solver(t_span=(0, t_end), t_steps=201, device="cuda", cache_dir=None)
    t, y = solver.integrate(ode, y0.to(cuda_device))
    assert y.shape == (201, 64, len(lo)) and y.dtype == dtype
    bits = 32 if dtype == torch.float32 else 64
    ref, st, _ = O.solve(sysid, y0.double().numpy(), _oracle_params(ode), t.double().cpu().numpy(),
                         0.0, bits=bits)
    assert (st == 0).all()
    got = y.double().cpu().numpy()
    scale = 1.0 + np.abs(ref)
    err_ic = (np.abs(got - ref) / scale).max(axis=(0, 2))
    if bits == 64:
        # within 20x the local error tolerance (atol 1e-6 + rtol 1e-8; the
        # same step rules with rounding-level differences in dt, amplified by
        # the Lorenz flow)
        bound = np.full(len(err_ic), 20 * (1e-6 + 1e-8))
    else:
        # float32: a few float32 ulps of the state times the steps, plus a
        # multiple of how far a one-ulp change of the float32 IC moves the
        # float64 oracle's own answer (ICs near a separatrix amplify rounding)
        p = _oracle_params(ode)
        ts = t.double().cpu().numpy()
        ref64, _, _ = O.solve(sysid, y0.double().numpy(), p, ts, 0.0, bits=64)
        sens = np.zeros(len(err_ic))
        for direction in (np.inf, -np.inf):
            yp = np.nextafter(y0.numpy(), np.float32(direction)).astype(np.float64)
            rp, _, _ = O.solve(sysid, yp, p, ts, 0.0, bits=64)
            sens = np.maximum(sens, (np.abs(rp - ref64) / scale).max(axis=(0, 2)))
        bound = 2e-4 + 4 * sens
    print(f"[result] {name} {dtype} max scaled err {err_ic.max():.3e} (worst IC "
          f"{int(np.argmax(err_ic / bound))}: {err_ic.max():.2e} against bound "
          f"{bound[int(np.argmax(err_ic / bound))]:.2e})", flush=True)
    assert (err_ic <= bound).all()
    # save time t0 returns y0 exactly
    assert torch.equal(y[0].cpu(), y0)


def _f32_bound(sysid, y0, p, ts, ref, t0=0.0):
    """Per-IC bound of the short-horizon float32 comparison: 2e-4 of the
    scaled state plus four times how far a one-ulp change of the float32 IC
    moves the float64 oracle's answer."""
    scale = 1.0 + np.abs(ref)
    ref64, _, _ = O.solve(sysid, y0.double().numpy(), p, ts, t0, bits=64)
    sens = np.zeros(y0.shape[0])
    for direction in (np.inf, -np.inf):
        yp = np.nextafter(y0.numpy(), np.float32(direction)).astype(np.float64)
        rp, _, _ = O.solve(sysid, yp, p, ts, t0, bits=64)
        sens = np.maximum(sens, (np.abs(rp - ref64) / scale).max(axis=(0, 2)))
    return 2e-4 + 4 * sens


@pytest.mark.parametrize("name", ["pendulum", "duffing"])
def test_float_lanes_outside_fast_trig_range_are_redone(cuda_device, name):
    """float32 evaluates the pendulum's sine and Duffing's end-time sincos by
    branch-free forms valid for |argument| <= 2^17; a lane whose step leaves
    that range is integrated again with the general forms.

    Pendulum: every other lane starts at an angle near +-1e10 rad, where the
    branch-free reduction is wrong (its quadrant overflows) and float32 holds
    the angle fixed (one ulp is 1024 rad), so omega relaxes to
    (T - K sin theta0) / alpha with the true sine: those lanes match the
    float32 oracle to 1e-5, and the in-range lanes of the launch give the
    same bits as a launch without them.  Duffing: times from 1.4e5, every
    lane redone, against the float32 oracle; there float32 stage times are
    multiples of 2^-6, and a rounding-level difference in a step size moves a
    stage time by one of them (a forcing change of A 2^-6 = 3e-3), so the
    bound is 5e-2 of the scaled state, which still separates a wrong
    forcing phase (errors of order one)."""
    ode, sysid, lo, hi, _ = CASES[name]
    print(f"[case] fast trig range {name}", flush=True)
    p = _oracle_params(ode)
    y0 = torch.tensor(O.sample_uniform(64, lo, hi, seed=11), dtype=torch.float32)
    t0 = 0.0
    far = torch.arange(64) % 2 == 1
    if name == "pendulum":
        k = torch.arange(32, dtype=torch.float32)
        y0[far, 0] = torch.where(k % 2 == 0, 1.0, -1.0) * (1.0e10 + 1024.0 * k)
    else:
        t0 = 140_000.0
    solver = HipDopri5Solver(t_span=(t0, t0 + 10.0), t_steps=101, device="cuda", cache_dir=None)
    t, y = solver.integrate(ode, y0.to(cuda_device))
    ts = t.double().cpu().numpy()
    ref, st, _ = O.solve(sysid, y0.double().numpy(), p, ts, t0, bits=32)
    assert (st == 0).all()
    got = y.double().cpu().numpy()
    err_ic = (np.abs(got - ref) / (1.0 + np.abs(ref))).max(axis=(0, 2))
    if name == "pendulum":
        fm = far.numpy()
        assert (got[:, fm, 0] == y0.numpy()[fm, 0]).all()  # the angle stays put in float32
        bound = np.where(fm, 1e-5, _f32_bound(sysid, y0, p, ts, ref, t0))
    else:
        bound = np.full(len(err_ic), 5e-2)
    print(f"[result] {name} worst ratio {float((err_ic / bound).max()):.3f}", flush=True)
    assert (err_ic <= bound).all()
    if name == "pendulum":
        _, y_near = solver.integrate(ode, y0[~far].contiguous().to(cuda_device))
        assert torch.equal(y_near.cpu(), y[:, ~far.to(y.device)].cpu())


def test_pendulum_golden_ics_float64(cuda_device):
    with open(os.path.join(GOLD, "zigode_pendulum_golden.json")) as f:
        gold = json.load(f)
    ics = torch.tensor(gold["ics"], dtype=torch.float64)
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cuda", cache_dir=None)
    t, y = solver.integrate(PendulumODE(PEND), ics.to(cuda_device))
    got = y.cpu().numpy()
    ref, _, _ = O.solve(O.SYS_PENDULUM, ics.numpy(), _oracle_params(PendulumODE(PEND)),
                        t.cpu().numpy(), 0.0, bits=64)
    tol = gold["tolerances"]
    # against the oracle's per-lane run: the same algorithm, rounding-level drift
    assert np.abs(got - ref).max() < 1e-3
    # against the reference's published final angles (zigode per IC and
    # torchdiffeq shared-step) within the reference's own cross-solver bounds
    for key in ("zig", "torchdiffeq"):
        assert np.abs(got[-1, :, 0] - np.array(gold["final_theta"][key])).max() < tol["CROSS_SOLVER_ATOL"]
    np.testing.assert_allclose(got[-1, [0, 4], 0], math.asin(0.5), atol=tol["FIXED_POINT_ATOL"])
    np.testing.assert_allclose(got[-1, [1, 2, 3], 1], ref[-1, [1, 2, 3], 1],
                               atol=tol["LIMIT_CYCLE_OMEGA_ATOL"])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_features_equal_statistics_of_written_window(cuda_device, dtype):
    y0 = torch.tensor(O.sample_uniform(512, [-1.0, -0.5], [1.0, 1.0], seed=3), dtype=dtype)
    solver = HipDopri5Solver(t_span=(0, 200), t_steps=500, t_eval=(150.0, 200.0), device="cuda",
                             cache_dir=None)
    res = solver.integrate_fused(DuffingODE(DUFF), y0.to(cuda_device), time_steady=170.0,
                                 keep_trajectories=True)
    si = res.steady_idx
    assert si == O.filter_time_idx(res.t.cpu().numpy(), 170.0)
    w = res.y[si:].double().cpu().numpy()
    f = O.features(w)
    stats = res.stats.cpu().numpy()
    for j, key in enumerate(["mean", "standard_deviation", "minimum", "maximum"]):
        np.testing.assert_allclose(stats[:, :, j], f[key], rtol=1e-9, atol=1e-9, err_msg=key)
    assert torch.equal(res.final, res.y[-1])
    assert (res.status.cpu() == 0).all()
    assert (res.nsteps.cpu() > 0).all()


def test_lane_refill_is_deterministic(cuda_device):
    """The same ICs through 64 resident lanes (heavy refill) and through a
    full device give bitwise identical results."""
    y0 = torch.tensor(O.sample_uniform(1000, [-1.0, -0.5], [1.0, 1.0], seed=11)).to(cuda_device)
    a = HipDopri5Solver(t_span=(0, 100), t_steps=100, device="cuda", cache_dir=None)
    b = HipDopri5Solver(t_span=(0, 100), t_steps=100, device="cuda", cache_dir=None, max_lanes=64)
    ra = a.integrate_fused(DuffingODE(DUFF), y0, keep_trajectories=True)
    rb = b.integrate_fused(DuffingODE(DUFF), y0, keep_trajectories=True)
    assert torch.equal(ra.y, rb.y) and torch.equal(ra.nsteps, rb.nsteps)
    assert torch.equal(ra.stats, rb.stats)


def test_lorenz_stop_freeze_matches_masked_oracle(cuda_device):
    """Per-lane stop at |y|inf >= 200 with the frozen state filled in equals
    integrating the masked RHS to the end (lorenz_ode.py:58-64)."""
    y0 = torch.tensor(O.sample_uniform(128, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], seed=5))
    stop = HipDopri5Solver(t_span=(0, 60), t_steps=120, device="cuda", cache_dir=None,
                           stop_threshold=200.0)
    t, y = stop.integrate(LorenzODE(LOR), y0.to(cuda_device))
    res = stop.integrate_fused(LorenzODE(LOR), y0.to(cuda_device), keep_trajectories=True)
    status = res.status.cpu().numpy()
    assert (status == _lib.BH_LANE_DIVERGED).any() and (status == 0).any()
    ref, _, _ = O.solve(O.SYS_LORENZ, y0.double().numpy(), _oracle_params(LorenzODE(LOR)),
                        t.double().cpu().numpy(), 0.0, bits=32)
    got = y.double().cpu().numpy()
    div = status == _lib.BH_LANE_DIVERGED
    # diverged lanes: the masked run has reached the freeze threshold too (the
    # stage-level mask can leave it a few ulps below 200 while it creeps up)
    assert (np.abs(ref[-1, div]).max(axis=1) >= 200 * (1 - 1e-5)).all()
    agree = np.abs(got[-1] - ref[-1]).max(axis=1) <= 1e-3 * (1 + np.abs(ref[-1]).max(axis=1))
    assert agree.mean() > 0.95
    # Inf-fill variant (JaxSolver Event semantics)
    inf = HipDopri5Solver(t_span=(0, 60), t_steps=120, device="cuda", cache_dir=None,
                          stop_threshold=200.0, inf_fill=True)
    r2 = inf.integrate_fused(LorenzODE(LOR), y0.to(cuda_device), keep_trajectories=True)
    yy = r2.y.cpu()
    d2 = r2.status.cpu().numpy() == _lib.BH_LANE_DIVERGED
    assert torch.isinf(yy[-1, torch.from_numpy(d2)]).all()
    assert torch.isfinite(yy[:, torch.from_numpy(~d2)]).all()


def test_status_max_steps_and_nonfinite(cuda_device):
    y0 = torch.tensor([[0.3, 0.2], [1.0, 0.0]], device=cuda_device)
    s = HipDopri5Solver(t_span=(0, 100), t_steps=10, device="cuda", cache_dir=None, max_steps=5)
    r = s.integrate_fused(PendulumODE(PEND), y0, keep_trajectories=True)
    assert (r.status.cpu() == _lib.BH_LANE_MAX_STEPS).all()
    assert (r.nsteps.cpu() == 5).all()
    bad = torch.tensor([[float("nan"), 0.0]], device=cuda_device)
    s2 = HipDopri5Solver(t_span=(0, 1), t_steps=5, t_eval=(0.5, 1.0), device="cuda", cache_dir=None)
    r2 = s2.integrate_fused(PendulumODE(PEND), bad, keep_trajectories=True)
    assert int(r2.status.cpu()[0]) == _lib.BH_LANE_NONFINITE


# ---- the reference's solver contract (tests/unit/solvers/test_solver.py) on the HIP path ----

def test_contract_decay(cuda_device):
    solver = HipDopri5Solver(t_span=(0, 1), t_steps=10, device="cuda", cache_dir=None)
    t, y = solver.integrate(ExponentialDecayODE({"decay": -1.0}), torch.tensor([[1.0]], device=cuda_device))
    assert t.shape == (10,) and y.shape == (10, 1, 1)
    assert y[0].item() == pytest.approx(1.0, abs=1e-5)
    assert y[-1].item() == pytest.approx(math.exp(-1), abs=1e-6)


def test_contract_t_eval_subrange(cuda_device):
    solver = HipDopri5Solver(t_span=(0, 2), t_steps=10, t_eval=(1, 2), device="cuda", cache_dir=None)
    t, y = solver.integrate(ExponentialDecayODE({"decay": -1.0}), torch.tensor([[1.0]], device=cuda_device))
    assert t[0].item() == pytest.approx(1.0) and t[-1].item() == pytest.approx(2.0)
    assert y[0].item() == pytest.approx(math.exp(-1), abs=1e-6)


@pytest.mark.parametrize("phase_iters", [0, 3])
def test_contract_params_ic_major(cuda_device, phase_iters):
    """IC-major parameter rows (base.py:137-159); also through the two-phase
    launch, whose resumed lanes keep their own parameter row."""
    solver = HipDopri5Solver(t_span=(0, 1), t_steps=11, device="cuda", cache_dir=None,
                             phase_iters=phase_iters)
    y0 = torch.tensor([[1.0], [2.0]], device=cuda_device)
    params = torch.tensor([[-1.0], [-2.0]])
    _, y = solver.integrate(ExponentialDecayODE({"decay": -1.0}), y0, params=params)
    assert y.shape == (11, 4, 1)
    want = [math.exp(-1), math.exp(-2), 2 * math.exp(-1), 2 * math.exp(-2)]
    np.testing.assert_allclose(y[-1, :, 0].cpu().numpy(), want, rtol=1e-5)
    # linearity: trajectory of y0=2 is twice that of y0=1 for each parameter
    np.testing.assert_allclose((y[:, 2] / y[:, 0]).cpu().numpy(), 2.0, rtol=1e-5)


def test_contract_errors(cuda_device):
    solver = HipDopri5Solver(t_span=(0, 1), t_steps=10, device="cuda", cache_dir=None)
    with pytest.raises(ValueError):
        solver.integrate(ExponentialDecayODE({"decay": -1.0}), torch.tensor([1.0], device=cuda_device))
    with pytest.raises(ValueError, match="t_eval"):
        HipDopri5Solver(t_span=(0, 1), t_steps=10, t_eval=(0.5, 1.5), device="cuda", cache_dir=None)
    c = HipDopri5Solver(t_span=(0, 2), t_eval=(1, 2), device="cuda", cache_dir=None).clone(device="cpu")
    assert c.t_eval == (1, 2) and c.device.type == "cpu"


@pytest.mark.parametrize("phase_iters", [0, 150])
def test_graph_replay_matches_eager(cuda_device, phase_iters):
    """The captured integration replayed several times, with eager work and
    synchronisation between replays, equals the eager launch every time
    (single launch, and the two-phase launch with its counter reset and
    compaction inside the graph)."""
    from pybasin_amd.graphs import CapturedIntegration

    ode = DuffingODE(DUFF)
    solver = HipDopri5Solver(t_span=(0, 100), t_steps=200, t_eval=(50.0, 100.0), device="cuda",
                             cache_dir=None, phase_iters=phase_iters)
    ya = torch.tensor(O.sample_uniform(3000, [-1.0, -0.5], [1.0, 1.0], seed=21)).to(cuda_device)
    yb = torch.tensor(O.sample_uniform(3000, [-1.0, -0.5], [1.0, 1.0], seed=22)).to(cuda_device)
    cap = CapturedIntegration(solver, ode, ya, time_steady=60.0)
    for y in (ya, yb, ya, yb):
        eager = solver.integrate_fused(ode, y, time_steady=60.0)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=cuda_device).sum()  # unrelated eager work between replays
        got = cap.replay(y)
        torch.cuda.synchronize()
        assert torch.equal(got.stats, eager.stats)
        assert torch.equal(got.final, eager.final) and torch.equal(got.nsteps, eager.nsteps)


def test_autotuner_picks_and_caches_a_lane_cap(cuda_device, tmp_path, monkeypatch):
    from pybasin_amd import autotune

    monkeypatch.setenv("PYBASIN_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    ode = DuffingODE(DUFF)
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=10, device="cuda", cache_dir=None)
    y0 = torch.tensor(O.sample_uniform(20000, [-1.0, -0.5], [1.0, 1.0], seed=2)).to(cuda_device)
    best = autotune.tune_lane_cap(solver, ode, y0, probe_span=20.0)
    assert best >= 0
    assert (tmp_path / "tune.json").exists()
    assert autotune.tune_lane_cap(solver, ode, y0, probe_span=20.0) == best  # cached
    cap, q = autotune.tune_lane_launch(solver, ode, y0, probe_span=20.0, caps=(0, 1), quorums=(64, 1))
    assert cap >= 0 and q in (64, 1)
    assert autotune.tune_lane_launch(solver, ode, y0, probe_span=20.0, caps=(0, 1),
                                     quorums=(64, 1)) == (cap, q)  # cached
    # with the two-phase launch among the candidates
    shape = autotune.tune_launch_shape(solver, ode, y0, probe_span=20.0, caps=(0,), quorums=(64,))
    assert set(shape) == {"max_lanes", "refill_quorum", "phase_iters"}
    assert any(k.startswith("phase") for k in autotune.last_timings)
    assert autotune.tune_launch_shape(solver, ode, y0, probe_span=20.0, caps=(0,),
                                      quorums=(64,)) == shape  # cached
    autotune.apply_launch_shape(solver, shape)
    assert solver.phase_iters == shape["phase_iters"]


def test_native_autotuner_solves_and_picks_the_fastest(cuda_device):
    """bh_autotune_lane: every (cap, quorum) candidate is timed, the fastest is
    returned, and the outputs it leaves equal a plain bh_solve_batch launch
    (the results do not depend on the launch shape)."""
    y0 = torch.tensor(O.sample_uniform(3000, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], seed=6),
                      dtype=torch.float32).to(cuda_device)
    ode = LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6})
    s = HipDopri5Solver(t_span=(0, 100), t_steps=200, t_eval=(50.0, 100.0), device="cuda",
                        cache_dir=None, stop_threshold=200.0, inf_fill=True)
    save_ts, y0p = s._prepare_tensors(y0)
    plan = s.plan_fused(ode, y0p, save_ts, None)
    caps, quorums, phases = [0, 128, 512], [1, 64], [20, 200]
    cap, q, k, timings, res = s.tune_fused(plan, caps, quorums, phases, reps=1)
    assert set(timings) == ({("refill", c, j) for c in caps for j in quorums}
                            | {("phased", j) for j in phases})
    assert all(t > 0 for t in timings.values())
    won = ("phased", k) if k else ("refill", cap, q)
    assert timings[won] == min(timings.values())
    base = s.launch_fused(plan, keep_trajectories=False)
    assert torch.equal(res.stats.nan_to_num(7.0), base.stats.nan_to_num(7.0))
    assert torch.equal(res.final.nan_to_num(7.0), base.final.nan_to_num(7.0))
    assert torch.equal(res.status, base.status) and torch.equal(res.nsteps, base.nsteps)
    assert (base.status == 4).any() and (base.status == 0).any()
    with pytest.raises(ValueError, match="quorum"):
        s.tune_fused(plan, [0], [65])
    # only phase candidates
    cap, q, k, timings, res = s.tune_fused(plan, [], [], [50], reps=1)
    assert (cap, q, k) == (0, 0, 50) and torch.equal(res.nsteps, base.nsteps)


@pytest.mark.parametrize("name,dtype,kw,K", [
    ("duffing", torch.float32, {}, 300), ("duffing", torch.float64, {}, 1),
    ("pendulum", torch.float32, {}, 250),
    ("lorenz", torch.float32, {"stop_threshold": 200.0}, 150),
    ("lorenz", torch.float32, {"stop_threshold": 200.0, "inf_fill": True}, 40),
    ("duffing", torch.float32, {}, 10**6),  # budget beyond every trajectory: nothing suspended
])
def test_two_phase_launch_equals_one_launch(cuda_device, name, dtype, kw, K):
    """bh_solve_batch_phased (phase 1 capped at K attempted steps, the rest
    packed and resumed) gives the trajectories, statistics and counters of a
    single launch bit for bit, on a ragged batch size."""
    ode, _, lo, hi, _ = CASES[name]
    n = 1000 + 37
    y0 = torch.tensor(O.sample_uniform(n, lo, hi, seed=11), dtype=dtype).to(cuda_device)
    common = dict(t_span=(0, 60), t_steps=120, t_eval=(30.0, 60.0), device="cuda", cache_dir=None, **kw)
    one = HipDopri5Solver(**common).integrate_fused(ode, y0, keep_trajectories=True)
    two = HipDopri5Solver(**common, phase_iters=K).integrate_fused(ode, y0, keep_trajectories=True)
    assert torch.equal(two.y.nan_to_num(7.0), one.y.nan_to_num(7.0))
    assert torch.equal(two.stats.nan_to_num(7.0), one.stats.nan_to_num(7.0))
    assert torch.equal(two.final.nan_to_num(7.0), one.final.nan_to_num(7.0))
    assert torch.equal(two.status, one.status) and torch.equal(two.nsteps, one.nsteps)
    if K < 10**6:
        assert int((one.nsteps > K).sum()) > 0  # some trajectories were suspended and resumed


def test_refill_quorum_does_not_change_results(cuda_device):
    """Per-lane and whole-wave refills (quorum 1 and 64) with divergence
    stops give bitwise identical fused results."""
    y0 = torch.tensor(O.sample_uniform(3000, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], seed=5),
                      dtype=torch.float32).to(cuda_device)
    ode = LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6})
    out = []
    for q in (1, 7, 64):
        s = HipDopri5Solver(t_span=(0, 100), t_steps=200, device="cuda", cache_dir=None,
                            stop_threshold=200.0, inf_fill=True, max_lanes=256, refill_quorum=q)
        out.append(s.integrate_fused(ode, y0, keep_trajectories=True))
    for r in out[1:]:
        assert torch.equal(r.y.nan_to_num(7.0), out[0].y.nan_to_num(7.0))
        assert torch.equal(r.stats.nan_to_num(7.0), out[0].stats.nan_to_num(7.0))
        assert torch.equal(r.status, out[0].status) and torch.equal(r.nsteps, out[0].nsteps)
    assert (out[0].status == 4).any() and (out[0].status == 0).any()
    with pytest.raises(ValueError):
        HipDopri5Solver(refill_quorum=65)


@pytest.mark.parametrize("name,dtype,kw", [
    ("duffing", torch.float32, {}), ("duffing", torch.float64, {}), ("pendulum", torch.float32, {}),
    ("lorenz", torch.float32, {"stop_threshold": 200.0}),
    ("lorenz", torch.float32, {"stop_threshold": 200.0, "inf_fill": True}),
    # refilled lanes (a resident set smaller than the batch)
    ("duffing", torch.float32, {"max_lanes": 64}),
    ("lorenz", torch.float32, {"stop_threshold": 200.0, "inf_fill": True, "max_lanes": 128,
                               "refill_quorum": 1}),
])
def test_fused_orbit_peaks_equal_trajectory_peaks(cuda_device, name, dtype, kw):
    """bh_solve_batch_peaks: the kernel's peaks equal extract_orbit_data on the
    trajectory the same launch wrote (same saved values, so bit for bit), and
    the other outputs equal a launch without peaks."""
    from pybasin_amd.orbit import extract_orbit_data, orbit_data_from_result

    ode, _, lo, hi, _ = CASES[name]
    y0 = torch.tensor(O.sample_uniform(700, lo, hi, 31), dtype=dtype).to(cuda_device)
    solver = HipDopri5Solver(t_span=(0, 400), t_steps=800, t_eval=(300.0, 400.0), device="cuda",
                             cache_dir=None, **kw)
    save_ts, y0p = solver._prepare_tensors(y0)
    plan = solver.plan_fused(ode, y0p, save_ts, None, time_steady=320.0)
    res = solver.launch_fused(plan, keep_trajectories=True, peak_cap=399)
    base = solver.launch_fused(plan, keep_trajectories=True)
    torch.cuda.synchronize()
    assert torch.equal(res.y, base.y) and torch.equal(res.stats.nan_to_num(7.0), base.stats.nan_to_num(7.0))
    dof = list(range(y0.shape[1]))
    got = orbit_data_from_result(res, dof, 320.0)
    want = extract_orbit_data(res.t, res.y, dof, time_steady=320.0)
    assert torch.equal(got.peak_counts, want.peak_counts)
    assert got.peak_values.shape == want.peak_values.shape
    assert torch.equal(got.peak_values.nan_to_num(-7.0), want.peak_values.nan_to_num(-7.0))
    assert int(got.peak_counts.max()) > 0
    print(f"[result] {name} {dtype}: max peaks {int(got.peak_counts.max())}", flush=True)


def test_orbit_peak_buffer_overflow_reruns(cuda_device):
    """A first buffer smaller than the peak count triggers one exact re-run."""
    from pybasin_amd.orbit import extract_orbit_data, integrate_orbit

    y0 = torch.tensor(O.sample_uniform(300, [-1.0, -0.5], [1.0, 1.0], 8)).to(cuda_device)
    solver = HipDopri5Solver(t_span=(0, 300), t_steps=600, t_eval=(200.0, 300.0), device="cuda",
                             cache_dir=None)
    res, od = integrate_orbit(solver, DuffingODE(DUFF), y0, dof=[0], time_steady=220.0, peak_cap=2)
    assert res.peak_values.shape[0] == int(res.peak_counts.max()) > 2
    full = solver.integrate_fused(DuffingODE(DUFF), y0, time_steady=220.0, keep_trajectories=True)
    want = extract_orbit_data(full.t, full.y, [0], time_steady=220.0)
    assert torch.equal(od.peak_counts, want.peak_counts)
    assert torch.equal(od.peak_values.nan_to_num(-7.0), want.peak_values.nan_to_num(-7.0))


def test_fused_results_cache_round_trip(cuda_device, tmp_path, monkeypatch):
    """integrate_fused with a cache_dir writes the fused outputs to safetensors;
    a second call with the same inputs reads them back without launching, a
    different feature window is a different entry, a corrupt file is
    recomputed (cache_manager.py:27-128 semantics)."""
    ode = DuffingODE(DUFF)
    solver = HipDopri5Solver(t_span=(0, 100), t_steps=200, t_eval=(50.0, 100.0), device="cuda",
                             cache_dir=str(tmp_path))
    y0 = torch.tensor(O.sample_uniform(500, [-1.0, -0.5], [1.0, 1.0], seed=3)).to(cuda_device)
    first = solver.integrate_fused(ode, y0, time_steady=60.0, keep_trajectories=True)
    files = list(tmp_path.iterdir())
    assert len(files) == 1
    calls = []
    real = HipDopri5Solver.run_fused
    monkeypatch.setattr(HipDopri5Solver, "run_fused",
                        lambda self, *a, **k: calls.append(1) or real(self, *a, **k))
    again = solver.integrate_fused(ode, y0, time_steady=60.0, keep_trajectories=True)
    assert not calls
    for f in ("t", "y", "final", "stats", "status", "nsteps"):
        assert torch.equal(getattr(again, f).cpu(), getattr(first, f).cpu()), f
    assert again.steady_idx == first.steady_idx and again.final.device == first.final.device
    solver.integrate_fused(ode, y0, time_steady=70.0, keep_trajectories=True)
    assert len(calls) == 1 and len(list(tmp_path.iterdir())) == 2
    files[0].write_bytes(b"corrupt")
    fixed = solver.integrate_fused(ode, y0, time_steady=60.0, keep_trajectories=True)
    assert len(calls) == 2 and torch.equal(fixed.stats.cpu(), first.stats.cpu())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_order_stats_match_torch(cuda_device, dtype):
    """bh_window_order_stats = torch's median / quantile order statistics on
    windows with ties, signed zeros, infinities and NaN."""
    from pybasin_amd.features import window_feature, window_order_stats

    g = torch.Generator().manual_seed(5)
    for N in (1, 2, 9, 1000):
        y = torch.randint(-50, 50, (N, 300, 3), generator=g).to(dtype) / 7
        y[:, 0, 0] = 0.0
        y[0, 0, 0] = -0.0
        y[N // 2, 1, 1] = float("inf")
        y[0, 2, 2] = -float("inf")
        y[N - 1, 3, 0] = float("nan")
        yd = y.to(cuda_device)
        start = N // 3
        n = N - start
        ranks = [0, (n - 1) // 2, n - 1]
        got = window_order_stats(yd, start, ranks).cpu()
        srt = torch.sort(y[start:], dim=0).values
        want = srt[ranks]
        nan = torch.isnan(y[start:]).any(0)
        want = torch.where(nan[None], torch.full_like(want, float("nan")), want)
        assert torch.equal(got.nan_to_num(123.0), want.nan_to_num(123.0)), N
        med = window_feature("median", None, yd, start).cpu()
        assert torch.equal(med.nan_to_num(123.0), y[start:].median(dim=0).values.nan_to_num(123.0))
        q = window_feature("quantile", {"q": 0.3}, yd, start).cpu()
        assert torch.allclose(q.nan_to_num(123.0), torch.quantile(y[start:], 0.3, dim=0).nan_to_num(123.0),
                              rtol=1e-6, atol=1e-6, equal_nan=True)


def test_minimal_feature_set_on_fused_result(cuda_device):
    """The reference's default feature set (median, mean, std, var, rms, max,
    abs max, min, delta, log_delta) from a fused launch with kept
    trajectories equals the torch calculators on that trajectory window."""
    from pybasin_amd.features import MINIMAL_FEATURES, HipFusedFeatureExtractor

    ode = DuffingODE(DUFF)
    solver = HipDopri5Solver(t_span=(0, 200), t_steps=400, t_eval=(100.0, 200.0), device="cuda",
                             cache_dir=None)
    y0 = torch.tensor(O.sample_uniform(2000, [-1.0, -0.5], [1.0, 1.0], seed=9)).to(cuda_device)
    res = solver.integrate_fused(ode, y0, time_steady=150.0, keep_trajectories=True)
    fx = HipFusedFeatureExtractor(time_steady=150.0, features=MINIMAL_FEATURES)
    f = fx.extract_from_result(res).cpu()
    w = res.y[res.steady_idx:].double().cpu()
    ref = {"median": w.median(dim=0).values, "mean": w.mean(0),
           "standard_deviation": w.std(0, correction=0), "variance": w.var(0, correction=0),
           "root_mean_square": torch.sqrt((w ** 2).mean(0)), "maximum": w.max(0).values,
           "absolute_maximum": w.abs().max(0).values, "minimum": w.min(0).values,
           "delta": (w.max(0).values - w.mean(0)).abs(),
           "log_delta": torch.log((w.max(0).values - w.mean(0)).abs() + 1e-10)}
    for j, name in enumerate(fx.feature_names):
        s, feat = int(name.split("__")[0][6:]), name.split("__")[1]
        np.testing.assert_allclose(f[:, j].numpy(), ref[feat][:, s].to(torch.float32).numpy(),
                                   rtol=2e-6, atol=2e-6, err_msg=name)
    with pytest.raises(ValueError, match="keep_trajectories"):
        fx.extract_from_result(solver.integrate_fused(ode, y0, time_steady=150.0))
