#!/usr/bin/env python3
"""Benchmark: classified initial conditions per second on one MI355X.

Default workload (the metric's own configuration): the pendulum torque sweep
through the learned surrogate with Dopri5 fallback.  46 torques T uniformly
in [0.1, 1.0] with per-torque counts np.random.default_rng(0).integers(
10_000, 200_001, 46) (4,431,800 ICs, theta in [-pi + asin T, pi + asin T],
omega in [-10, 10]), alpha 0.1, K 1, t in [0, 1000], rtol 1e-8, atol 1e-6,
float32; labels FP / LC from the window [950, 1000].  The network (trained
once, outside the timed steps, by the project's kernels on 50,000
integrator-labelled ICs) labels confident ICs; the rest go to the per-lane
Dopri5 kernel; a 2 % audit per torque is integrated over the full span.  A
step is one gated pass over all 4,431,800 ICs; `value` is the pass rate with
four passes in flight (each on its own stream and host thread, taking the
next pass when it has issued one, as a server would overlap the
latency-bound tail of one pass with the next; four measured best against
two, six and eight), and the plain path (full
integration of every IC, with the same passes in flight and its own tuned
launch shape) is timed in the same run; the basin-stability error of the
gated labels against the plain labels is measured, not assumed.  The plain
Duffing workload (100,000 ICs, float32) is measured as well and reported in
`config.duffing_100k`.

Other workloads: --workload duffing | pendulum_c1 | lorenz | rossler |
mlp_infer | hdbscan | estimator | estimator_defaults.  Inputs are resident in HBM before timing
starts.  Prints one JSON line.  --gpus above 1 is not supported: this build
targets one GPU and runs no collectives.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VALU = 157.3e12  # MI355X FP32 vector peak (packed), FLOP/s
PEAK_FP64_VALU = 78.6e12
PEAK_BF16_MFMA = 2500e12   # dense

# flops per attempted step: 6 F_rhs + 63 S + 16 (RK combination, error norm,
# controller); a step whose span holds a save point adds the dense-output fit
# (36 S); each save point adds its evaluation (9 S) and the streaming
# statistics (6 S) when it lies in the feature window
F_RHS = {"duffing": 8, "pendulum": 5, "lorenz": 9}
STATE = {"duffing": 2, "pendulum": 2, "lorenz": 3}
MLP_MACS = 6 * 256 + 256 * 256 * 2 + 256 * 2  # multiply-adds per row (unpadded layers)

# HBM bytes per dispatch of the Duffing 100k float32 kernel, from rocprofv3
# FETCH_SIZE + WRITE_SIZE scaled by a plain-copy calibration
# (profiles/dopri5_lane_duffing_100k_f32_pmc.md)
TRAFFIC_BYTES = {("duffing", "f32", 100_000, False): 9.50e6}
# the two-phase launch adds the round trip of each suspended trajectory's
# state record and index through HBM, per suspended trajectory (same
# counters at the tuner's phase budget 2,779, 64,422 suspended: 3.725e7
# bytes per launch pair; profiles/dopri5_phase_duffing_100k_f32.md)
TRAFFIC_PER_SUSPENDED = {("duffing", "f32", 100_000): 431.0}

DUFFING_TEMPLATES = [[-0.21, 0.02], [1.05, 0.77], [-0.67, 0.02], [-0.46, 0.30], [-0.43, 0.12]]


def _label_counts(labels: torch.Tensor, k: int) -> torch.Tensor:
    """Counts per label 0..k-1 on the device without a host round trip
    (torch.bincount reads the maximum label back to size its output)."""
    return torch.zeros(k, dtype=torch.int64, device=labels.device).scatter_add_(
        0, labels.long(), torch.ones_like(labels, dtype=torch.int64))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-lanes", type=int, default=0)
    ap.add_argument("--phase-iters", type=int, default=0,
                    help="duffing, lorenz: two-phase launch with this step budget instead of the tuner")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape tuner (every trajectory on its own lane, one launch)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--keep-traj", action="store_true",
                    help="duffing: also write the full trajectories to HBM (reports write GB/s)")
    ap.add_argument("--workload", choices=["surrogate", "duffing", "pendulum_c1", "lorenz", "rossler",
                                           "mlp_infer", "window_features", "window_spectrum",
                                           "comprehensive",
                                           "hdbscan", "estimator", "estimator_defaults"],
                    default="surrogate")
    ap.add_argument("--gate", type=float, default=0.999,
                    help="surrogate: confidence gate of the headline pass")
    ap.add_argument("--config-gate", type=float, default=0.95,
                    help="surrogate: the configured gate, run and reported beside the headline")
    ap.add_argument("--inflight", type=int, default=4,
                    help="surrogate: passes in flight (each on its own stream and host thread)")
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--rossler-coupling", choices=["edges", "dense"], default="edges",
                    help="rossler workload: the edge-list wavefront kernel or the dense MFMA tile kernel")
    ap.add_argument("--evolve-t", type=float, default=50.0,
                    help="surrogate: below-gate ICs are integrated to this time and classified "
                         "again before a full integration (0 disables)")
    ap.add_argument("--no-duffing", action="store_true",
                    help="surrogate: skip the plain Duffing measurement in config.duffing_100k")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration():
    n = 64 * 1024 * 1024  # 64 Mi float32 = 256 MiB
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 3}), flush=True)


# ---- host cores and the CPU baselines ------------------------------------------

def host_cores() -> dict:
    """Cores this process may use, and what os.cpu_count() reports.  Thread
    and process pools use the cores the process is allowed to run on, at most
    16 (the GPU box grants one GPU's job a 16-CPU share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"os_cpu_count": os.cpu_count(), "affinity": avail, "used": max(1, min(avail, 16))}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _rhs_scipy(sysname, p):
    if sysname == "duffing":
        d, k3, A = p
        return lambda t, y: [y[1], -d * y[1] - k3 * y[0] ** 3 + A * math.cos(t)]
    if sysname == "pendulum":
        a, T, K = p
        return lambda t, y: [y[1], -a * y[1] + T - K * math.sin(y[0])]
    s, r, b = p

    def f(t, y):
        if max(abs(y[0]), abs(y[1]), abs(y[2])) >= 200.0:  # lorenz_ode.py:58-64 freeze mask
            return [0.0, 0.0, 0.0]
        return [s * (y[1] - y[0]), r * y[0] - y[0] * y[2] - y[1], y[0] * y[1] - b * y[2]]
    return f


def _scipy_one(sysname, y0, p, save, t1, rtol, atol):
    from scipy.integrate import solve_ivp

    r = solve_ivp(_rhs_scipy(sysname, p), (0.0, t1), y0, method="RK45", rtol=rtol, atol=atol,
                  max_step=t1 / 100.0, t_eval=save)
    return r.y[:, -1]


def cpu_baselines(sysname, y0_cpu, params_rows, save_ts, *, n_torch=4096, n_scipy=64, n_oracle=20_000,
                  rtol=1e-8, atol=1e-6, t1=1000.0):
    """The reference's CPU paths on seeded samples of the workload's own rows
    (same float32 ICs and per-row parameters), timed on the host cores.

    * TorchDiffEqSolver counterpart (torchdiffeq_solver.py:97-134): one
      shared-step Dopri5 over the sample in torch on the CPU (this package's
      TorchDopri5Solver, torchdiffeq's algorithm), float32, the whole span
      with the workload's save grid -- the cpu_baseline.
    * ScipyParallelSolver counterpart (scipy_solver.py:113-169): solve_ivp
      RK45 per IC over a process pool, max_step = (t1 - t0) / 100, float64.
    * The oracle: per-lane Dopri5 in C over threads (the same algorithm as
      the kernel), float32.
    Returns (cpu_baseline, [the other two])."""
    import multiprocessing as mp

    import numpy as np

    from oracle import oracle as O
    from pybasin_amd.solvers import TorchDopri5Solver
    from pybasin_amd.systems import DuffingODE, LorenzODE, PendulumODE

    cores = host_cores()
    used = cores["used"]
    ode = {"duffing": DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}),
           "pendulum": PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}),
           "lorenz": LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6})}[sysname]
    sid = {"duffing": O.SYS_DUFFING, "pendulum": O.SYS_PENDULUM, "lorenz": O.SYS_LORENZ}[sysname]
    save32 = save_ts.to(torch.float32).cpu()
    save64 = save32.double().numpy()
    others = []
    # ScipyParallelSolver counterpart (before torch spins up its thread pool)
    ys = y0_cpu[:n_scipy].double().numpy()
    ps = params_rows[:n_scipy].double().numpy()
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(used) as pool:  # the GPU is not initialised in this process yet
        pool.starmap(_scipy_one, [(sysname, y, p, save64, t1, rtol, atol) for y, p in zip(ys, ps)])
    dt = time.perf_counter() - t0
    others.append({"path": "ScipyParallelSolver counterpart (solve_ivp RK45 per IC, process pool)",
                   "value": round(n_scipy / dt, 1), "unit": "ICs/s", "cores": used,
                   "sample": f"first {n_scipy} rows, float64, t in [0,{t1:g}], {dt:.2f} s"})
    # TorchDiffEqSolver counterpart
    torch.set_num_threads(used)
    solver = TorchDopri5Solver(t_span=(0, t1), t_steps=len(save32), device="cpu", rtol=rtol, atol=atol,
                               cache_dir=None)
    yt = y0_cpu[:n_torch].to(torch.float32).contiguous()
    pt = params_rows[:n_torch].to(torch.float32).contiguous()
    t0 = time.perf_counter()
    solver._integrate(ode, yt, save32, pt)
    dt = time.perf_counter() - t0
    base = {"value": round(n_torch / dt, 1), "unit": "ICs/s", "cores": used, "kind": "port",
            "path": "TorchDiffEqSolver counterpart (shared-step Dopri5, torch CPU)",
            "sample": f"first {n_torch} rows of the workload (seeded), float32, t in [0,{t1:g}], "
                      f"{len(save32)} saves, one shared-step batch: {dt:.2f} s",
            "cpu_model": cpu_model(), "os_cpu_count": cores["os_cpu_count"],
            "affinity_cores": cores["affinity"]}
    # the oracle, per lane
    yo = y0_cpu[:n_oracle].double().numpy()
    po = params_rows[:n_oracle].double().numpy()
    t0 = time.perf_counter()
    O.solve(sid, yo, po, save64, 0.0, rtol=rtol, atol=atol, bits=32, threads=used, t1=None)
    dt = time.perf_counter() - t0
    others.append({"path": "oracle (per-lane Dopri5 in C, pthreads; the kernel's algorithm)",
                   "value": round(len(yo) / dt, 1), "unit": "ICs/s", "cores": used,
                   "sample": f"first {len(yo)} rows, float32, {dt:.2f} s"})
    return base, others


# ---- passes in flight -------------------------------------------------------------

def run_passes(fn, k: int, inflight: int):
    """k calls of fn with up to `inflight` in flight: `inflight` host threads,
    each issuing on a stream of its own, take the next of the k calls as they
    finish issuing the previous one.  Returns (wall ms per call between two
    device-wide synchronisations, the last result of the first thread)."""
    streams = [torch.cuda.Stream() for _ in range(inflight)]
    out = [None] * inflight
    err = []
    lock = threading.Lock()
    left = [k]

    def take():
        with lock:
            if left[0] == 0:
                return False
            left[0] -= 1
            return True

    def worker(i):
        try:
            with torch.cuda.stream(streams[i]):
                while take():
                    out[i] = fn()
        except BaseException as e:  # surfaced after the join
            err.append(e)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if inflight == 1:
        worker(0)
    else:
        th = [threading.Thread(target=worker, args=(i,)) for i in range(inflight)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / k
    if err:
        raise err[0]
    return ms, out[0]


def _lane_flops(sysname, nsteps, save_hits, n_saved, window_saves):
    """Algorithmic FLOP of a per-lane launch: attempted steps x F_step, steps
    whose span holds a save x the dense-output fit (36 S), evaluated saves x
    9 S, window saves x the statistics (6 S)."""
    S = STATE[sysname]
    return (float(nsteps) * (6 * F_RHS[sysname] + 63 * S + 16) + float(save_hits) * 36 * S
            + float(n_saved) * 9 * S + float(window_saves) * 6 * S)


def surrogate_workload(args):
    """The metric's configuration (see the module docstring)."""
    import numpy as np

    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.surrogate import (basin_stability_by_group, classify_gated, integrate_labels,
                                       torque_sweep_ics, train_surrogate, tune_plain_launch)
    from pybasin_amd.systems import PendulumODE

    T = np.linspace(0.1, 1.0, 46)
    counts = np.random.default_rng(args.seed).integers(10_000, 200_001, 46)
    x, group = torque_sweep_ics(T, counts, seed=args.seed, device="cpu")
    n = x.shape[0]
    save_ts = torch.linspace(900.0, 1000.0, 1000, dtype=torch.float32)
    base, others = None, None
    if not args.no_cpu:
        perm = torch.randperm(n, generator=torch.Generator().manual_seed(args.seed + 7))
        xs = x[perm]
        prow = torch.stack([torch.full((n,), 0.1), xs[:, 2], torch.full((n,), 1.0)], 1)
        base, others = cpu_baselines("pendulum", xs[:, :2], prow, save_ts, n_torch=4096)
        base["sample"] = "seeded random rows of the sweep (all torques): " + base["sample"]
    x, group = x.to("cuda"), group.to("cuda")
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             rtol=1e-8, atol=1e-6, cache_dir=None)
    ode = PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0})
    ts = 950.0
    ev = args.evolve_t if args.evolve_t > 0 else None
    # one-time training on 50,000 integrator-labelled ICs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model, hist, _ = train_surrogate(solver, ode, x, n_train=50_000, epochs=args.epochs, seed=args.seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # the plain path: its default launch and its tuned launch shape
    tsolver = solver.clone(cache_dir=None)
    t0 = time.perf_counter()
    launch = tune_plain_launch(tsolver, ode, x, ts) if not args.no_tune else {}
    launch["tune_s"] = round(time.perf_counter() - t0, 2)
    plain = {}
    for name, sv in (("default", solver), ("tuned", tsolver)):
        run_passes(lambda: integrate_labels(sv, ode, x, ts), max(1, args.warmup), 1)
        for fl in sorted({1, args.inflight}):
            plain[(name, fl)], lab = run_passes(lambda: integrate_labels(sv, ode, x, ts), args.steps, fl)
        plain[(name, "labels")] = lab
    plab = plain[("default", "labels")]
    bs_p = basin_stability_by_group(plab, group, 46)

    def gated(gate, stage_times=None):
        return classify_gated(model, solver, ode, x, gate=gate, group=group, time_steady=ts,
                              evolve_t=ev, n_groups=46, stage_times=stage_times)

    gated_ms = {}
    for gate in (args.gate, args.config_gate):
        run_passes(lambda: gated(gate), max(1, args.warmup), 1)
        for fl in sorted({1, args.inflight}):
            gated_ms[(gate, fl)], res = run_passes(lambda: gated(gate), args.steps, fl)
        gated_ms[(gate, "res")] = res
    res = gated_ms[(args.gate, "res")]
    res_cfg = gated_ms[(args.config_gate, "res")]
    err = (basin_stability_by_group(res.labels, group, 46) - bs_p).abs()
    err_cfg = (basin_stability_by_group(res_cfg.labels, group, 46) - bs_p).abs()
    audit_err = float((res.audit_surrogate != res.audit_integrated).float().mean())
    # one instrumented pass: stage times, the audit launch and the network pass
    stages = {}
    gated(args.gate, stages)
    audit_steps = stages.pop("audit_nsteps")
    audit_rows = stages.pop("audit_rows")
    audit_ms = stages.pop("audit_launch_ms")
    infer_ms = stages.pop("infer_launch_ms", None)
    # the audit launch's algorithmic FLOP: its rows' attempted and save-hit
    # step counts (the same trajectories re-solved with the count flag)
    from pybasin_amd.solvers.hip_dopri5 import count_save_steps

    xa = x[res.audit_idx]
    pa = torch.stack([torch.full((len(xa),), 0.1, dtype=torch.float64, device=x.device), xa[:, 2].double(),
                      torch.full((len(xa),), 1.0, dtype=torch.float64, device=x.device)], 1).contiguous()
    sa, ya = solver._prepare_tensors(xa[:, :2].contiguous())
    a_steps, a_hits = count_save_steps(solver, ode, ya, sa, pa, ts)
    assert a_steps == int(audit_steps.sum()) and len(xa) == audit_rows
    from pybasin_amd.solvers.hip_dopri5 import steady_index

    win = 1000 - steady_index(sa, ts)
    audit_flops = _lane_flops("pendulum", a_steps, a_hits, audit_rows * 1000, audit_rows * win)
    # the plain pass's Dopri5 launch (tuned shape) timed on its stream, with
    # save-hit step counts from the kernel
    plain_roof = plain_launch_roofline(tsolver, ode, x, ts)
    ms_g = gated_ms[(args.gate, args.inflight)]
    ms_tuned = plain[("tuned", args.inflight)]
    out = {
        "metric": "classified initial conditions/sec at <=0.5% basin-stability error vs plain",
        "value": round(n / (ms_g * 1e-3), 1), "unit": "ICs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_g, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
        "data": "synthetic: seeded-uniform ICs per torque value; network trained on the "
                "integrator's labels of 50,000 of them",
        "config": {
            "workload": "pendulum_torque_sweep_surrogate", "n_ics": n, "torques": 46,
            "counts": "np.random.default_rng(0).integers(10_000, 200_001, 46)",
            "t_span": [0, 1000], "rtol": 1e-8, "atol": 1e-6, "label_window": [950, 1000],
            "gate": args.gate, "inflight": args.inflight,
            "max_bs_error_vs_plain": round(float(err.max()), 5),
            "worst_torque": round(float(T[int(err.argmax())]), 3),
            "ms_per_pass_one_in_flight": round(gated_ms[(args.gate, 1)], 3),
            "plain_ms": {"default": round(plain[("default", args.inflight)], 3),
                         "tuned": round(ms_tuned, 3),
                         "default_one_in_flight": round(plain[("default", 1)], 3),
                         "tuned_one_in_flight": round(plain[("tuned", 1)], 3)},
            "plain_tuned_launch": launch,
            "plain_tuned_labels_equal_default": bool(torch.equal(plain[("tuned", "labels")], plab)),
            "speedup_vs_plain_tuned": round(ms_tuned / ms_g, 3),
            "speedup_vs_plain_default": round(plain[("default", args.inflight)] / ms_g, 3),
            "speedup_one_in_flight_vs_plain_tuned": round(plain[("tuned", 1)] / gated_ms[(args.gate, 1)], 3),
            "routed_frac": round(float(res.routed.float().mean()), 4),
            "evolve_t": ev, "evolved_frac": round(float(res.evolved.float().mean()), 4),
            "integrated_frac": round(float(res.integrated.float().mean()), 4),
            "late_integrated": res.n_late,
            "audit_n": len(res.audit_idx), "audit_network_label_error": round(audit_err, 5),
            "stage_ms": {k: round(v, 3) for k, v in stages.items()},
            "configured_gate": {
                "gate": args.config_gate,
                "max_bs_error_vs_plain": round(float(err_cfg.max()), 5),
                "worst_torque": round(float(T[int(err_cfg.argmax())]), 3),
                "within_0.5pct": bool(float(err_cfg.max()) <= 0.005),
                "ms_per_pass": round(gated_ms[(args.config_gate, args.inflight)], 3),
                "ms_per_pass_one_in_flight": round(gated_ms[(args.config_gate, 1)], 3),
                "routed_frac": round(float(res_cfg.routed.float().mean()), 4)},
            "train_n": 50_000, "train_epochs": args.epochs, "train_s": round(t_train, 2),
            "final_train_loss": round(hist[-1], 4),
        },
        "roofline": {
            "bound": "valu", "kernel": "bh::dopri5_lane_kernel<Pendulum,float> (the audit launch: "
                                       "the gated pass's longest kernel, on its critical path)",
            "achieved": round(audit_flops / (audit_ms * 1e-3) / 1e12, 3), "peak": 157.3,
            "unit": "TFLOP/s", "frac": round(audit_flops / (audit_ms * 1e-3) / PEAK_FP32_VALU, 4),
            "traffic": None, "kernel_ms": round(audit_ms, 3), "rows": audit_rows,
            "attempted_steps": a_steps, "save_hit_steps": a_hits,
            "mean_steps": round(float(audit_steps.double().mean()), 1),
            "max_steps": int(audit_steps.max())},
        "roofline_plain": plain_roof,
    }
    if infer_ms is not None:
        fl = 2.0 * MLP_MACS * n
        out["roofline_network"] = {
            "bound": "mfma", "kernel": "bh::mlp::mlp_infer_kernel (beside the audit)",
            "achieved": round(fl / (infer_ms * 1e-3) / 1e12, 1), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(fl / (infer_ms * 1e-3) / PEAK_BF16_MFMA, 4), "kernel_ms": round(infer_ms, 3)}
    if base is not None:
        out["cpu_baseline"] = base
        out["cpu_other_paths"] = others
    if not args.no_duffing:
        d = duffing_workload(args, cpu=False, emit=False)
        out["config"]["duffing_100k"] = {
            "value": d["value"], "unit": "ICs/s", "ms_per_step": d["ms_per_step"],
            "launch": d["config"]["launch"], "roofline": d["roofline"]}
    print(json.dumps(out), flush=True)


def plain_launch_roofline(solver, ode, x, time_steady, reps: int = 3):
    """The plain pass's Dopri5 launch (the solver's launch shape) on all rows,
    timed with HIP events on its stream, FLOP from the kernel's per-lane
    attempted and save-hit step counts."""
    from pybasin_amd.solvers.hip_dopri5 import count_save_steps

    n = x.shape[0]
    alpha, K = float(ode.params["alpha"]), float(ode.params["K"])
    params = torch.stack([torch.full((n,), alpha, dtype=torch.float64, device=x.device),
                          x[:, 2].double(), torch.full((n,), K, dtype=torch.float64, device=x.device)],
                         1).contiguous()
    save_ts, y0 = solver._prepare_tensors(x[:, :2].contiguous())
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = solver.run_fused(ode, y0, save_ts, params, keep_trajectories=False,
                               time_steady=time_steady, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    steps, hits = count_save_steps(solver, ode, y0, save_ts, params, time_steady)
    window = 1000 - int(res.steady_idx)
    ok = int((res.status == 0).sum())
    flops = _lane_flops("pendulum", steps, hits, ok * 1000, ok * window)
    k = sum(ms) / len(ms)
    kern = (f"bh::dopri5_phase_kernel<Pendulum,float> phases 1 + 2 (K = {solver.phase_iters})"
            if solver.phase_iters else "bh::dopri5_lane_kernel<Pendulum,float>")
    return {"bound": "valu", "kernel": kern, "achieved": round(flops / (k * 1e-3) / 1e12, 3),
            "peak": 157.3, "unit": "TFLOP/s", "frac": round(flops / (k * 1e-3) / PEAK_FP32_VALU, 4),
            "traffic": None, "kernel_ms": round(k, 3), "rows": n,
            "attempted_steps": int(steps), "save_hit_steps": int(hits)}


def rossler_cpu_baseline(y0_cpu, Ks, n_ics: int = 100):
    """The oracle (per-lane Dopri5 in C with the edge scatter-add coupling of
    rossler_network_jax_ode.py:72-73 and the stop at |y|inf >= 400) on the
    first n_ics ICs x all coupling strengths, float32.  The reference runs
    this case with JaxSolver; a shared-step torch run cannot, since its
    unbounded trajectories end the whole batch."""
    import numpy as np

    from oracle import oracle as O
    from pybasin_amd.systems import rossler_edges

    cores = host_cores()
    O.set_rossler_edges(rossler_edges())
    rows = np.repeat(y0_cpu[:n_ics].double().numpy(), len(Ks), axis=0)
    prow = np.tile(np.array([[0.2, 0.2, 7.0, k] for k in Ks]), (n_ics, 1))
    t0 = time.perf_counter()
    O.solve(O.SYS_ROSSLER, rows, prow, np.array([1000.0]), 0.0, rtol=1e-3, atol=1e-6, bits=32,
            threads=cores["used"])
    dt = time.perf_counter() - t0
    return {"value": round(len(rows) / dt, 1), "unit": "ICs/s", "cores": cores["used"], "kind": "port",
            "path": "oracle (per-lane Dopri5 in C, edge scatter-add coupling, stop at 400)",
            "sample": f"first {n_ics} ICs x {len(Ks)} coupling strengths, float32, t in [0,1000], "
                      f"{dt:.2f} s", "cpu_model": cpu_model(), "os_cpu_count": cores["os_cpu_count"],
            "affinity_cores": cores["affinity"]}


def rossler_workload(args):
    """Rossler network, N = 100 nodes on the fixed Watts-Strogatz graph, 11
    coupling strengths K in [0.119, 0.159] x 10,000 ICs each (IC-major, one
    launch), t in [0, 1000], rtol 1e-3, atol 1e-6, stop at |y|inf >= 400;
    labels synchronized / desynchronized / unbounded from the final state."""
    import numpy as np

    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.solvers.rossler import sync_labels
    from pybasin_amd.systems import RosslerNetworkODE

    B = min(args.n, 10_000)
    Ks = np.linspace(0.119, 0.159, 11)
    torch.manual_seed(args.seed)
    lo = torch.tensor([-15.0] * 200 + [-5.0] * 100)
    hi = torch.tensor([15.0] * 200 + [35.0] * 100)
    y0_cpu = torch.rand(B, 300) * (hi - lo) + lo
    base = None
    if not args.no_cpu:
        base = rossler_cpu_baseline(y0_cpu, Ks)
    y0 = y0_cpu.to("cuda")
    params = torch.tensor([[0.2, 0.2, 7.0, k] for k in Ks], dtype=torch.float64)
    ode = RosslerNetworkODE({"a": 0.2, "b": 0.2, "c": 7.0, "K": 0.119})
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=2, t_eval=(999.0, 1000.0), device="cuda",
                             rtol=1e-3, atol=1e-6, cache_dir=None, stop_threshold=400.0)
    solver.rossler_coupling = args.rossler_coupling
    stream = torch.cuda.current_stream()
    kev = []

    def step(record=False):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = solver.integrate_fused(ode, y0, params)
        e1.record(stream)
        if record:
            kev.append((e0, e1))
        return res

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step(record=True)
    lab = sync_labels(res.final, res.status)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    kernel_ms = sum(a.elapsed_time(b) for a, b in kev) / len(kev)
    # the other coupling form once, outside the timed steps, for comparison
    other = "dense" if args.rossler_coupling == "edges" else "edges"
    solver.rossler_coupling = other
    solver.integrate_fused(ode, y0, params)  # warm-up
    o0, o1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    o0.record(stream)
    res_o = solver.integrate_fused(ode, y0, params)
    o1.record(stream)
    torch.cuda.synchronize()
    other_ms = o0.elapsed_time(o1)
    lab_o = sync_labels(res_o.final, res_o.status)
    n = B * len(Ks)
    nsteps = res.nsteps.double()
    # FLOP per attempted step: 6 evaluations x (coupling + 8 N local) + 63 S
    # RK arithmetic + 16, the coupling being the reference's edge form (2 FLOP
    # per edge entry: x_i - x_j and its scatter-add; 800 entries) for the
    # wavefront kernel and the 2 N^2 product for the dense matrix-core kernel
    steps_sum = float(nsteps.sum())
    rk_flop = steps_sum * (6 * 8 * 100 + 63 * 300 + 16)
    if args.rossler_coupling == "edges":
        flop = rk_flop + steps_sum * 6 * 2 * 800
        roof = {"bound": "valu", "achieved": round(flop / (kernel_ms * 1e-3) / 1e12, 3),
                "peak": 157.3, "unit": "TFLOP/s", "frac": round(flop / (kernel_ms * 1e-3) / 157.3e12, 4),
                "traffic": None, "kernel": "bh::ross::wave_kernel (edge-list coupling)",
                "kernel_ms": round(kernel_ms, 3)}
    else:
        mfma_flop = steps_sum * 6 * 2 * 100 * 100
        roof = {"bound": "mfma", "achieved": round(mfma_flop / (kernel_ms * 1e-3) / 1e12, 3),
                "peak": 157.3, "unit": "TFLOP/s",
                "frac": round(mfma_flop / (kernel_ms * 1e-3) / 157.3e12, 4), "traffic": None,
                "valu_tflops": round(rk_flop / (kernel_ms * 1e-3) / 1e12, 3),
                "kernel": "bh::ross::tile_kernel (dense matrix-core coupling)",
                "kernel_ms": round(kernel_ms, 3)}
    per_k = {}
    for i, k in enumerate(Ks):
        li = lab[i::len(Ks)]
        per_k[f"{k:.3f}"] = {c: round(float(np.mean(li == c)), 4) for c in ("synchronized", "desynchronized", "unbounded")}
    out = {
        "metric": "classified initial conditions/sec at <=0.5% basin-stability error vs plain",
        "value": round(n / (ms * 1e-3), 1), "unit": "ICs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded-uniform ICs, 11 coupling strengths in one launch",
        "config": {"workload": "rossler100_k_sweep", "coupling": args.rossler_coupling,
                   "n_trajectories": n, "ics_per_k": B,
                   "rtol": 1e-3, "atol": 1e-6, "t_span": [0, 1000],
                   "mean_steps": round(float(nsteps.mean()), 1), "max_steps": int(nsteps.max()),
                   "basin_stability": per_k,
                   "other_coupling": {"coupling": other, "kernel_ms": round(other_ms, 3),
                                      "label_agreement": round(float(np.mean(lab == lab_o)), 4)}},
        "roofline": roof,
    }
    if base is not None:
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


def mlp_infer_workload(args):
    """The fused surrogate inference kernel alone on the torque-sweep rows
    (4.43M (theta, omega, T) rows, random-init weights): rows/s and MFMA
    roofline (133,120 multiply-adds per row, bf16 dense peak 2.5 PFLOP/s)."""
    import numpy as np

    from pybasin_amd.surrogate import SurrogateMLP, torque_sweep_ics

    T = np.linspace(0.1, 1.0, 46)
    counts = np.random.default_rng(args.seed).integers(10_000, 200_001, 46)
    x, _ = torque_sweep_ics(T, counts, seed=args.seed)
    n = x.shape[0]
    model = SurrogateMLP(device="cuda:0", seed=args.seed)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        model.predict(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        model.predict(x)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    kernel_ms = e0.elapsed_time(e1) / args.steps
    macs = 6 * 256 + 256 * 256 * 2 + 256 * 2  # multiply-adds per row (unpadded layers)
    flop = 2.0 * macs * n
    out = {
        "metric": "surrogate inference rows/sec", "value": round(n / (ms * 1e-3), 1),
        "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic: torque-sweep rows, random-init weights",
        "config": {"workload": "mlp_infer_torque_sweep", "rows": n, "net": "6(3)-256-256-256-2 SiLU"},
        "roofline": {"bound": "mfma", "achieved": round(flop / (kernel_ms * 1e-3) / 1e12, 1),
                     "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(flop / (kernel_ms * 1e-3) / 2500e12, 4), "traffic": None,
                     "kernel": "bh::mlp::mlp_infer_kernel", "kernel_ms": round(kernel_ms, 3)},
    }
    print(json.dumps(out), flush=True)


def _duffing_window(args, dev):
    """The kept window of the Duffing workload: n = args.n seeded-uniform ICs
    (+ 5 templates), 1,000 saves on [900, 1000], integrated once (not timed)."""
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import DuffingODE

    dtype = torch.float32 if args.dtype == "f32" else torch.float64
    torch.manual_seed(args.seed)
    lo, hi = torch.tensor([-1.0, -0.5]), torch.tensor([1.0, 1.0])
    y0 = torch.cat([torch.rand(args.n, 2) * (hi - lo) + lo,
                    torch.tensor(DUFFING_TEMPLATES)]).to(dtype).to(dev)
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             rtol=1e-8, atol=1e-6, cache_dir=None)
    res = solver.integrate_fused(DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), y0,
                                 time_steady=900.0, keep_trajectories=True)
    return res.y


def window_spectrum_workload(args):
    """The spectrum calculators of the reference's comprehensive set
    (fft_coefficient k = 0..99 x real / imag / abs / angle, fft_aggregated,
    spkt_welch_density, fourier_entropy: 412 features per state) over the kept
    Duffing window (n = 100,000 ICs + 5 templates, float32, 1,000 rows); a step
    = bh_window_spectrum (bins 0..99 and the aggregates of all 501 bins) plus
    the named features on the device.  Roofline: the kernel is bound by its
    double-precision vector issue: 3 flops (one FMA, one subtraction) per bin
    and sample, all 501 bins of each series (the aggregates need them), over
    the 78.6 TFLOP/s FP64 vector peak."""
    from pybasin_amd.window_features import SPECTRUM_FEATURES, spectrum_feature, window_spectrum

    dev = torch.device("cuda:0")
    y = _duffing_window(args, dev)
    N, B, S = y.shape
    plan = [(name, p) for name, grid in SPECTRUM_FEATURES.items() for p in grid]
    stream = torch.cuda.current_stream(dev)

    def step():
        c, a = window_spectrum(y, 0, 100, True)
        return [spectrum_feature(name, p, c, a, N) for name, p in plan]

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    k0.record(stream)
    for _ in range(args.steps):
        window_spectrum(y, 0, 100, True)
    k1.record(stream)
    torch.cuda.synchronize()
    kernel_ms = k0.elapsed_time(k1) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    nb = N // 2 + 1
    flops = 3.0 * nb * N * B * S
    out = {
        "metric": "window spectrum features, series/sec", "value": round(B * S / (ms * 1e-3), 1),
        "unit": "series/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: kept Duffing windows of seeded-uniform ICs",
        "config": {"workload": "window_spectrum_duffing", "n_ics": B, "window_rows": N, "states": S,
                   "series_dtype": args.dtype, "features_per_state": len(plan),
                   "kernel": "bh::win::spectrum_kernel"},
        "roofline": {"bound": "fp64-valu", "achieved": round(flops / (kernel_ms * 1e-3) / 1e12, 2),
                     "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(flops / (kernel_ms * 1e-3) / 78.6e12, 4), "traffic": None,
                     "kernel_ms": round(kernel_ms, 3), "algorithmic_flops": flops},
    }
    print(json.dumps(out), flush=True)


def comprehensive_workload(args):
    """Every calculator of the reference's comprehensive set built here
    (settings.py:223-340) over the kept Duffing window (n = 100,000 ICs + 5
    templates, float32, 1,000 rows, 2 states), integrated once outside the
    timed region; a step = all window kernels (window-feature slots,
    spectrum, sort + change quantiles, autocovariances, chunk trends / fits /
    wavelets, peak counts, ordinal patterns, the delay-embedding estimates,
    the Lempel-Ziv parses) plus every named feature of every state.  Each
    kernel is also timed alone with HIP events.  Roofline: HBM over the
    kernels' window reads (block 2, spectrum 1, sort 1, change quantiles 2,
    autocovariances 1, trends 2, peaks 1, patterns 1, embeddings 1 each,
    Lempel-Ziv 1) and their float64 outputs, against the summed kernel time,
    are reported beside the roofline of the dominant kernel (Eckmann's
    exponents: FP64 issue of its pairwise Chebyshev passes); the O(n^2)
    embedding estimates and the sequential parses are bound by their own
    issue, not by HBM."""
    from pybasin_amd.features import window_feature
    from pybasin_amd.window_features import (COMPREHENSIVE_FEATURES, WindowBlocks, window_autocov,
                                             window_embedding, window_feature_block,
                                             window_lempel_ziv, window_peak_counts,
                                             window_perm_entropy, window_sorted, window_spectrum,
                                             window_trends)

    dev = torch.device("cuda:0")
    y = _duffing_window(args, dev)
    N, B, S = y.shape
    names = [(name, p) for name, grid in COMPREHENSIVE_FEATURES.items() for p in (grid or [None])]
    plan = [(0, name, p) for name, p in names]

    def step():
        wb = WindowBlocks(y, 0, plan, median=lambda: window_feature("median", None, y, 0))
        return [wb.feature(name, p) for name, p in names]

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    kernels = {
        "window_features": (lambda: window_feature_block(y, 0), 2, 100),
        "spectrum": (lambda: window_spectrum(y, 0, 100, True), 1, 205),
        "sorted+change_quantiles": (lambda: window_sorted(y, 0, True), 3, 71),
        "autocov": (lambda: window_autocov(y, 0, 40), 1, 41),
        "trends": (lambda: window_trends(y, 0, True), 2, 114),
        "peak_counts": (lambda: window_peak_counts(y, 0, [10, 50]), 1, 12),
        "perm_entropy": (lambda: window_perm_entropy(y, 0), 1, 5),
        "corr_dim": (lambda: window_embedding(y, 0, "corr_dim"), 1, 1),
        "lyap_r": (lambda: window_embedding(y, 0, "lyap_r"), 1, 1),
        "lyap_e": (lambda: window_embedding(y, 0, "lyap_e"), 1, 4),
        "lempel_ziv": (lambda: window_lempel_ziv(y, 0, [2, 3, 5, 10, 100]), 1, 5),
    }
    kernel_ms, alg_bytes = {}, 0.0
    for name, (fn, reads, rows) in kernels.items():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        kernel_ms[name] = round(e0.elapsed_time(e1) / args.steps, 3)
        alg_bytes += reads * N * B * S * y.element_size() + rows * 8.0 * B * S
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    ksum = sum(kernel_ms.values())
    # the dominant kernel (Eckmann's exponents): two Chebyshev passes over the
    # point pairs, a subtraction and a maximum per coordinate, in double
    no = N - 12
    le_flops = 2.0 * no * (no - 1) * 10 * 2 * B * S
    le_ms = kernel_ms["lyap_e"]
    out = {
        "metric": "comprehensive window features, series/sec", "value": round(B * S / (ms * 1e-3), 1),
        "unit": "series/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.dtype, "data": "synthetic: kept Duffing windows of seeded-uniform ICs",
        "config": {"workload": "comprehensive_duffing", "n_ics": B, "window_rows": N, "states": S,
                   "calculators": len(COMPREHENSIVE_FEATURES), "features_per_state": len(names)},
        "kernels_ms": kernel_ms,
        "roofline": {"bound": "fp64-valu", "kernel": "bh::win::lyap_e_kernel",
                     "achieved": round(le_flops / (le_ms * 1e-3) / 1e12, 2), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(le_flops / (le_ms * 1e-3) / 78.6e12, 4),
                     "traffic": None, "kernel_ms": le_ms, "algorithmic_flops": le_flops},
        "window_reads": {"bound": "hbm", "kernels_ms": round(ksum, 3), "algorithmic_bytes": alg_bytes,
                         "achieved_gbs": round(alg_bytes / (ksum * 1e-3) / 1e9, 1)},
    }
    print(json.dumps(out), flush=True)


def window_features_workload(args):
    """The comprehensive window features of the reference's comprehensive set
    (bh_window_features, then the named features on the host) over the kept
    window of the Duffing workload: n = 100,000 ICs (+ 5 templates), float32,
    1,000 saves on [900, 1000], integrated once outside the timed region; a
    step = the kernel's two passes over the (1000, n, 2) window plus the 128
    named features per state.  Roofline: HBM, two reads of the window
    (algorithmic bytes) plus the 90 float64 slots written per series."""
    from pybasin_amd.window_features import (COMPREHENSIVE_FEATURES, comprehensive_feature,
                                             feature_family, window_feature_block)

    dev = torch.device("cuda:0")
    y = _duffing_window(args, dev)
    N, B, S = y.shape
    plan = [(name, p) for name, grid in COMPREHENSIVE_FEATURES.items() for p in (grid or [None])
            if name != "symmetry_looking" and feature_family(name, p) == "block"]
    stream = torch.cuda.current_stream(dev)

    def step():
        blk = window_feature_block(y, 0)
        memo = {}
        return [comprehensive_feature(name, p, blk, N, cache=memo) for name, p in plan]

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    k0.record(stream)
    for _ in range(args.steps):
        window_feature_block(y, 0)
    k1.record(stream)
    torch.cuda.synchronize()
    kernel_ms = k0.elapsed_time(k1) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    elem = y.element_size()
    from pybasin_amd._lib import BH_WF_SLOTS

    alg_bytes = 2.0 * N * B * S * elem + BH_WF_SLOTS * 8.0 * B * S
    out = {
        "metric": "comprehensive window features, series/sec", "value": round(B * S / (ms * 1e-3), 1),
        "unit": "series/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.dtype, "data": "synthetic: kept Duffing windows of seeded-uniform ICs",
        "config": {"workload": "window_features_duffing", "n_ics": B, "window_rows": N, "states": S,
                   "features_per_state": len(plan), "kernel": "bh::win::window_features_kernel"},
        "roofline": {"bound": "hbm", "achieved": round(alg_bytes / (kernel_ms * 1e-3) / 1e9, 1),
                     "peak": 8000.0, "unit": "GB/s",
                     "frac": round(alg_bytes / (kernel_ms * 1e-3) / 8e12, 4), "traffic": None,
                     "kernel_ms": round(kernel_ms, 3), "algorithmic_bytes": alg_bytes},
    }
    print(json.dumps(out), flush=True)


def hdbscan_workload(args):
    """The reference's default predictor (HDBSCANClusterer(auto_tune=True,
    assign_noise=True), hdbscan_clusterer.py:12-148) on the GPU: Duffing
    features (max x, std x) of the n default ICs, integrated once outside the
    timed region; a step = one auto-tuned clustering with noise assignment.
    sklearn's own pipeline is timed on a bounded sample for comparison."""
    from pybasin_amd.cluster import HipHDBSCANClusterer
    from pybasin_amd.features import feature_from_stats
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import DuffingODE

    n = args.n
    torch.manual_seed(args.seed)
    lo = torch.tensor([-1.0, -0.5], dtype=torch.float64)
    hi = torch.tensor([1.0, 1.0], dtype=torch.float64)
    y0 = (torch.rand(n, 2, dtype=torch.float64) * (hi - lo) + lo).to("cuda")
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             cache_dir=None)
    res = solver.integrate_fused(DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), y0,
                                 time_steady=900.0)
    X = torch.stack([feature_from_stats("maximum", res.stats)[:, 0],
                     feature_from_stats("standard_deviation", res.stats)[:, 0]], 1)
    X = X.to(torch.float32).double().cpu().numpy()
    for _ in range(args.warmup):
        HipHDBSCANClusterer(auto_tune=True, assign_noise=True).fit_predict(X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c = HipHDBSCANClusterer(auto_tune=True, assign_noise=True)
        lab = c.fit_predict(X)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    out = {
        "metric": "clustered feature rows/sec (auto-tuned HDBSCAN + noise assignment)",
        "value": round(n / (ms * 1e-3), 1), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Duffing (max x, std x) features of seeded-uniform ICs",
        "config": {"workload": "hdbscan_auto_tune_duffing_features", "rows": n,
                   "chosen_min_cluster_size": c.min_cluster_size,
                   "clusters": int(len(set(lab.tolist()))),
                   "silhouette_scores": {str(k): round(v, 4) for k, v in c.scores_.items()}},
    }
    if not args.no_cpu:
        from sklearn.cluster import HDBSCAN
        from sklearn.metrics import silhouette_score

        m = min(n, 10_000)
        Xs = X[:m]
        t0 = time.perf_counter()
        sizes = [max(10, int(0.005 * m)), max(25, int(0.01 * m)), max(50, int(0.02 * m)),
                 max(100, int(0.03 * m)), max(150, int(0.05 * m))]
        for size in sizes:
            lb = HDBSCAN(min_cluster_size=size, min_samples=min(10, size // 5)).fit_predict(Xs)
            keep = lb != -1
            if len(set(lb[keep].tolist())) >= 2:
                silhouette_score(Xs[keep], lb[keep])
        HDBSCAN(min_cluster_size=sizes[0], min_samples=min(10, sizes[0] // 5)).fit_predict(Xs)
        t_cpu = time.perf_counter() - t0
        t0 = time.perf_counter()
        HipHDBSCANClusterer(auto_tune=True, assign_noise=True).fit_predict(Xs)
        torch.cuda.synchronize()
        t_gpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / t_cpu, 1), "unit": "rows/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"first {m} rows; sklearn {__import__('sklearn').__version__} "
                                         f"HDBSCAN x 6 + silhouette_score x 5 (the reference's "
                                         f"auto-tune steps): {t_cpu:.2f} s; this GPU path on the "
                                         f"same sample {t_gpu_s:.3f} s"}
    print(json.dumps(out), flush=True)


def estimator_workload(args):
    """The reference's end-to-end benchmark
    (benchmarks/end_to_end/benchmark_basin_stability_estimator.py): the
    pendulum (alpha 0.1, T 0.5, K 1) through BasinStabilityEstimator with
    n = 100,000 sampled ICs, t in [0, 1000] with 1,000 saves over the whole
    span, rtol 1e-8, atol 1e-6, PendulumFeatureExtractor(time_steady=950),
    1-NN on the templates (0.5, 0) FP and (2.7, 0) LC, no unbounded
    detection.  A step = one estimator run (sampling, integration, features,
    classification, basin stability) on the GPU."""
    import math

    from pybasin_amd.case_features import PendulumFeatureExtractor
    from pybasin_amd.estimator import BasinStabilityEstimator, UniformRandomSampler
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import PendulumODE

    n = args.n
    T = 0.5
    sampler = UniformRandomSampler([-math.pi + math.asin(T), -10.0], [math.pi + math.asin(T), 10.0],
                                   device="cuda")
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cuda", rtol=1e-8, atol=1e-6,
                             cache_dir=None)
    bse = BasinStabilityEstimator(PendulumODE({"alpha": 0.1, "T": T, "K": 1.0}), sampler, n=n,
                                  solver=solver,
                                  feature_extractor=PendulumFeatureExtractor(time_steady=950.0),
                                  template_y0=[[0.5, 0.0], [2.7, 0.0]], template_labels=["FP", "LC"],
                                  detect_unbounded=False)
    torch.manual_seed(args.seed)
    for _ in range(args.warmup):
        bse.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = bse.run()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    published = 100_000 / 12.5678  # ICs/s, RTX 5070 Ti Laptop, JAX solver
    value = n / (ms * 1e-3)
    out = {
        "metric": "classified initial conditions/sec (end-to-end BasinStabilityEstimator)",
        "value": round(value, 1), "unit": "ICs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / published, 2) if n == 100_000 else None,
        "dtype": "f32", "data": "synthetic: seeded-uniform ICs (the estimator's sampler on the GPU)",
        "config": {"workload": "pendulum_end_to_end_estimator", "n_ics": n,
                   "basin_stability": {k: round(v, 5) for k, v in res.basin_stability.items()},
                   "reference_published": {"value": round(published, 1), "unit": "ICs/s",
                                           "hardware": "RTX 5070 Ti Laptop, JAX/Diffrax solver",
                                           "source": "benchmarks/end_to_end/results/"
                                                     "end_to_end_comparison.csv:31"}},
    }
    print(json.dumps(out), flush=True)


def estimator_defaults_workload(args):
    """The reference's example run with pybasin's defaults
    (docs/benchmarks/basin-stability-estimator.md:9-24): the pendulum case
    study with n = 10,000 ICs where only the ODE system and the sampler
    region are given; solver, feature extractor (minimal set, z-scored) and
    predictor (auto-tuned HDBSCAN with noise assignment) take their
    defaults.  A step = one estimator run."""
    import math

    from pybasin_amd.estimator import BasinStabilityEstimator, UniformRandomSampler
    from pybasin_amd.systems import PendulumODE

    n = 10_000 if args.n == 100_000 else args.n
    T = 0.5
    sampler = UniformRandomSampler([-math.pi + math.asin(T), -10.0], [math.pi + math.asin(T), 10.0],
                                   device="cuda")
    torch.manual_seed(args.seed)
    bse = BasinStabilityEstimator(PendulumODE({"alpha": 0.1, "T": T, "K": 1.0}), sampler, n=n)
    for _ in range(args.warmup):
        bse.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = bse.run()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    value = n / (ms * 1e-3)
    published = 10_000 / 17.0062
    out = {
        "metric": "classified initial conditions/sec (end-to-end BasinStabilityEstimator, defaults)",
        "value": round(value, 1), "unit": "ICs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": round(value / published, 1) if n == 10_000 else None,
        "dtype": "f32", "data": "synthetic: seeded-uniform ICs (the estimator's sampler on the GPU)",
        "config": {"workload": "pendulum_estimator_defaults", "n_ics": n,
                   "basin_stability": {str(k): round(v, 5) for k, v in res.basin_stability.items()},
                   "reference_published": {"value": round(published, 1), "unit": "ICs/s",
                                           "total_s": 17.0062,
                                           "source": "docs/benchmarks/basin-stability-estimator.md:13-24"}},
    }
    print(json.dumps(out), flush=True)


def lorenz_workload(args):
    """Lorenz with escape: 200,000 seeded ICs (x in [-10, 10], y in [-20, 20],
    z = 0), sigma 0.12, r 0, b -0.6, t in [0, 1000], float32, rtol 1e-8,
    atol 1e-6, per-lane stop at |y|inf >= 200 with Inf fill, 1,000 saves on
    [900, 1000]; the launch shape (resident-lane cap and refill quorum, or
    the two-phase launch and its step budget) chosen by the autotuner; labels: unbounded by status, the two attractors by 1-NN on
    mean x (setup_lorenz_system.py:54-70)."""
    from pybasin_amd import autotune
    from pybasin_amd.cluster import knn
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import LorenzODE

    n = 200_000 if args.n == 100_000 else args.n
    torch.manual_seed(args.seed)
    lo = torch.tensor([-10.0, -20.0, 0.0])
    hi = torch.tensor([10.0, 20.0, 0.0])
    y0 = (torch.rand(n, 3) * (hi - lo) + lo)
    base = others = None
    if not args.no_cpu:
        prow = torch.tensor([[0.12, 0.0, -0.6]]).expand(n, 3)
        # the shared step follows the batch through every escape: ~10 ms a row
        base, others = cpu_baselines("lorenz", y0, prow, torch.linspace(900.0, 1000.0, 1000),
                                     n_torch=1024)
    templates = torch.tensor([[0.8, -3.0, 0.0], [-0.8, 3.0, 0.0]])
    y0_all = torch.cat([y0, templates]).to("cuda")
    ode = LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6})
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             rtol=1e-8, atol=1e-6, cache_dir=None, stop_threshold=200.0,
                             inf_fill=True)
    os.environ.setdefault("PYBASIN_AMD_TUNE_CACHE", os.path.join("/tmp", "pybasin_amd_tune.json"))
    t0 = time.perf_counter()
    if args.phase_iters > 0:
        shape = {"max_lanes": 0, "refill_quorum": 0, "phase_iters": args.phase_iters}
        autotune.last_timings.clear()
    else:
        shape = autotune.tune_launch_shape(solver, ode, y0_all, use_cache=False)
    t_tune = time.perf_counter() - t0
    autotune.apply_launch_shape(solver, shape)
    save_ts = torch.linspace(900.0, 1000.0, 1000, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    kev = []

    def step(record=False):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = solver.run_fused(ode, y0_all, save_ts, None, keep_trajectories=False, steady_idx=0,
                               stream=stream)
        e1.record(stream)
        if record:
            kev.append((e0, e1))
        mean_x = res.stats[:, 0, 0:1].contiguous()
        lab = knn(mean_x[:n], mean_x[n:], 1)[1][:, 0]
        lab = torch.where(res.status[:n] == 4, torch.full_like(lab, 2), lab)  # BH_LANE_DIVERGED
        return res, _label_counts(lab, 3)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, counts = step(record=True)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    kernel_ms = sum(a.elapsed_time(b) for a, b in kev) / len(kev)
    from pybasin_amd.solvers.hip_dopri5 import count_save_steps

    nsteps = res.nsteps.double()
    steps, hits = count_save_steps(solver, ode, y0_all, save_ts, None, None, steady_idx=0)
    ok = int((res.status == 0).sum())
    flops = _lane_flops("lorenz", steps, hits, ok * 1000, ok * 1000)
    out = {
        "metric": "classified initial conditions/sec at <=0.5% basin-stability error vs plain",
        "value": round(n / (ms * 1e-3), 1), "unit": "ICs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded-uniform ICs (torch.manual_seed, CPU float32 draw)",
        "config": {"workload": "lorenz_escape_hip_dopri5", "n_ics": n, "stop": "|y|inf >= 200, Inf fill",
                   "lane_cap": shape["max_lanes"], "refill_quorum": shape["refill_quorum"],
                   "phase_iters": shape["phase_iters"],
                   "tuner_timings_ms": autotune.last_timings,
                   "tune_s": round(t_tune, 2),
                   "basin_stability": {k: round(float(v) / n, 5) for k, v in
                                       zip(("attractor 1", "attractor 2", "unbounded"), counts.tolist())},
                   "mean_steps_per_ic": round(float(nsteps.mean()), 1),
                   "max_steps_per_ic": int(nsteps.max())},
        "roofline": {"bound": "valu", "achieved": round(flops / (kernel_ms * 1e-3) / 1e12, 3),
                     "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(flops / (kernel_ms * 1e-3) / PEAK_FP32_VALU, 4), "traffic": None,
                     "kernel": ("bh::dopri5_phase_kernel<Lorenz,float> phases 1 + 2" if solver.phase_iters
                                else "bh::dopri5_lane_kernel<Lorenz,float>"),
                     "kernel_ms": round(kernel_ms, 3),
                     "attempted_steps": int(steps), "save_hit_steps": int(hits)},
    }
    if base is not None:
        out["cpu_baseline"] = base
        out["cpu_other_paths"] = others
    print(json.dumps(out), flush=True)


def duffing_workload(args, cpu: bool = True, emit: bool = True):
    """Duffing oscillator x' = v, v' = -0.08 v - x^3 + 0.2 cos t, n = 100,000
    ICs seeded-uniform in x in [-1, 1], v in [-0.5, 1]; t in [0, 1000], rtol
    1e-8, atol 1e-6, float32 (--dtype f64 for float64); save grid 1,000
    points on [900, 1000]; fused mean / std / min / max per state over that
    window; 1-NN classification against the 5 template ICs of the
    reference's Duffing setup, integrated in the same launch.  The launch
    shape is chosen once by the project's autotuner before the timed steps."""
    from pybasin_amd.cluster import knn
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.solvers.hip_dopri5 import count_save_steps
    from pybasin_amd.systems import DuffingODE

    dev = torch.device("cuda:0")
    dtype = torch.float32 if args.dtype == "f32" else torch.float64
    n = args.n
    torch.manual_seed(args.seed)
    lo = torch.tensor([-1.0, -0.5])
    hi = torch.tensor([1.0, 1.0])
    y0_cpu = torch.rand(n, 2) * (hi - lo) + lo
    base = others = None
    save_cpu = torch.linspace(900.0, 1000.0, 1000, dtype=torch.float32)
    if cpu and not args.no_cpu:
        prow = torch.tensor([[0.08, 1.0, 0.2]]).expand(n, 3)
        base, others = cpu_baselines("duffing", y0_cpu, prow, save_cpu)
    y0_all = torch.cat([y0_cpu, torch.tensor(DUFFING_TEMPLATES)]).to(dtype).to(dev)
    ode = DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2})
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             rtol=1e-8, atol=1e-6, cache_dir=None, max_lanes=args.max_lanes)
    launch = {"max_lanes": args.max_lanes, "refill_quorum": 0, "phase_iters": args.phase_iters}
    solver.phase_iters = args.phase_iters
    if not args.no_tune and args.max_lanes == 0 and args.phase_iters == 0 and not args.keep_traj:
        # one-time launch-shape choice on this problem, outside the timed steps
        from pybasin_amd import autotune

        os.environ.setdefault("PYBASIN_AMD_TUNE_CACHE", os.path.join("/tmp", "pybasin_amd_tune.json"))
        t0 = time.perf_counter()
        launch = autotune.tune_launch_shape(solver, ode, y0_all, use_cache=False)
        launch = dict(launch, tune_s=round(time.perf_counter() - t0, 2),
                      tuner_timings_ms=dict(autotune.last_timings))
        autotune.apply_launch_shape(solver, launch)
    save_ts = torch.linspace(900.0, 1000.0, 1000, dtype=dtype, device=dev)
    stream = torch.cuda.current_stream(dev)
    kev = []

    def step(record=False):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = solver.run_fused(ode, y0_all, save_ts, None, keep_trajectories=args.keep_traj,
                               steady_idx=0, stream=stream)
        e1.record(stream)
        if record:
            kev.append((e0, e1))
        # reference Duffing classifier features: state 0 maximum and std; 1-NN
        # on the templates by the HIP k-NN kernel
        feats = torch.stack([res.stats[:, 0, 3], res.stats[:, 0, 1]], dim=1)
        labels = knn(feats[:n].contiguous(), feats[n:].contiguous(), 1)[1][:, 0]
        counts = _label_counts(labels, len(DUFFING_TEMPLATES))
        return res, counts

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, counts = step(record=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    kernel_ms = sum(a.elapsed_time(b) for a, b in kev) / len(kev)

    steps, hits = count_save_steps(solver, ode, y0_all, save_ts, None, None, steady_idx=0)
    status = res.status.cpu()
    ok = int((status == 0).sum())
    flops = _lane_flops("duffing", steps, hits, ok * 1000, ok * 1000)
    achieved = flops / (kernel_ms * 1e-3)
    peak = PEAK_FP32_VALU if dtype == torch.float32 else PEAK_FP64_VALU
    tname = "float" if dtype == torch.float32 else "double"
    bs = (counts.double() / n).cpu().tolist()
    nsteps = res.nsteps.to(torch.float64)
    traffic = None
    base_traffic = TRAFFIC_BYTES.get(("duffing", args.dtype, n, False))
    if not args.keep_traj and base_traffic is not None:
        if not solver.phase_iters:
            traffic = base_traffic
        elif TRAFFIC_PER_SUSPENDED.get(("duffing", args.dtype, n)) is not None:
            traffic = base_traffic + TRAFFIC_PER_SUSPENDED[("duffing", args.dtype, n)] * float(
                (nsteps > solver.phase_iters).sum())

    out = {
        "metric": "classified initial conditions/sec at <=0.5% basin-stability error vs plain",
        "value": round(n / (ms_step * 1e-3), 1),
        "unit": "ICs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: seeded-uniform ICs (torch.manual_seed, CPU float32 draw)",
        "config": {
            "workload": "duffing_plain_hip_dopri5",
            "n_ics": n, "t_span": [0, 1000], "save_window": [900, 1000], "t_steps": 1000,
            "rtol": 1e-8, "atol": 1e-6, "features": "fused mean/std/min/max per state",
            "classifier": "1-NN on (max x, std x), 5 templates",
            "basin_stability": [round(v, 5) for v in bs],
            "mean_steps_per_ic": round(float(nsteps.mean()), 1),
            "max_steps_per_ic": int(nsteps.max()),
            "lanes_failed": int((status != 0).sum()),
            "launch": dict(launch, suspended=int((nsteps > solver.phase_iters).sum())
                           if solver.phase_iters else 0),
        },
        "roofline": {
            "bound": "valu", "achieved": round(achieved / 1e12, 3),
            "peak": round(peak / 1e12, 1), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic,
            "kernel": (f"bh::dopri5_phase_kernel<Duffing,{tname}> phases 1 + 2" if solver.phase_iters
                       else f"bh::dopri5_lane_kernel<Duffing,{tname}>"),
            "kernel_ms": round(kernel_ms, 3),
            "attempted_steps": int(steps), "save_hit_steps": int(hits),
        },
    }
    if args.keep_traj:
        # trajectories (1000, n + 5, 2) written to HBM inside the kernel
        traj_bytes = float(res.y.numel() * res.y.element_size())
        out["config"]["trajectories"] = "kept (n_save, n, S) in HBM"
        out["traj_write"] = {"bytes": traj_bytes, "GBps": round(traj_bytes / (kernel_ms * 1e-3) / 1e9, 1),
                             "hbm_peak_GBps": 8000.0}
    if base is not None:
        out["cpu_baseline"] = base
        out["cpu_other_paths"] = others
    if emit:
        print(json.dumps(out), flush=True)
    return out


def pendulum_c1_workload(args):
    """The CPU-runnable configuration: forced damped pendulum (alpha 0.1,
    T 0.5, K 1), n = 2,000 ICs seeded-uniform in theta in [-pi + asin 0.5,
    pi + asin 0.5], omega in [-10, 10]; t in [0, 1000], rtol 1e-8, atol
    1e-6, float32; 1,000 saves on [900, 1000]; features per-state mean and
    std of that window (fused); 1-NN on the FP / LC templates (0.4, 0) and
    (2.7, 0) of setup_pendulum_system.py:53-59, integrated in the same
    launch.  The CPU paths run the whole workload (the scipy one on a
    sample)."""
    from pybasin_amd.cluster import knn
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import PendulumODE

    n = 2000
    a = math.asin(0.5)
    torch.manual_seed(args.seed)
    lo = torch.tensor([-math.pi + a, -10.0])
    hi = torch.tensor([math.pi + a, 10.0])
    y0 = torch.rand(n, 2) * (hi - lo) + lo
    save_cpu = torch.linspace(900.0, 1000.0, 1000, dtype=torch.float32)
    base = others = None
    if not args.no_cpu:
        prow = torch.tensor([[0.1, 0.5, 1.0]]).expand(n, 3)
        base, others = cpu_baselines("pendulum", y0, prow, save_cpu, n_torch=n, n_scipy=128,
                                     n_oracle=n)
        base["sample"] = "the whole workload: " + base["sample"]
    templates = torch.tensor([[0.4, 0.0], [2.7, 0.0]])
    y_all = torch.cat([y0, templates]).to("cuda")
    ode = PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0})
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             rtol=1e-8, atol=1e-6, cache_dir=None)
    save_ts = save_cpu.to("cuda")
    stream = torch.cuda.current_stream()
    kev = []

    def step(record=False):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = solver.run_fused(ode, y_all, save_ts, None, keep_trajectories=False, steady_idx=0,
                               stream=stream)
        e1.record(stream)
        if record:
            kev.append((e0, e1))
        feats = torch.stack([res.stats[:, 0, 0], res.stats[:, 0, 1], res.stats[:, 1, 0],
                             res.stats[:, 1, 1]], 1).contiguous()  # mean, std per state
        lab = knn(feats[:n].contiguous(), feats[n:].contiguous(), 1)[1][:, 0]
        return res, _label_counts(lab, 2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, counts = step(record=True)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    kernel_ms = sum(x.elapsed_time(y) for x, y in kev) / len(kev)
    from pybasin_amd.solvers.hip_dopri5 import count_save_steps

    steps, hits = count_save_steps(solver, ode, y_all, save_ts, None, None, steady_idx=0)
    flops = _lane_flops("pendulum", steps, hits, (n + 2) * 1000, (n + 2) * 1000)
    out = {
        "metric": "classified initial conditions/sec at <=0.5% basin-stability error vs plain",
        "value": round(n / (ms * 1e-3), 1), "unit": "ICs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded-uniform ICs (torch.manual_seed, CPU float32 draw)",
        "config": {"workload": "pendulum_c1_plain_hip_dopri5", "n_ics": n, "t_span": [0, 1000],
                   "save_window": [900, 1000], "features": "fused mean/std per state",
                   "classifier": "1-NN, templates (0.4, 0) FP / (2.7, 0) LC",
                   "basin_stability": {"FP": round(float(counts[0]) / n, 5),
                                       "LC": round(float(counts[1]) / n, 5)},
                   "max_steps_per_ic": int(res.nsteps.max())},
        "roofline": {"bound": "valu", "kernel": "bh::dopri5_lane_kernel<Pendulum,float>",
                     "achieved": round(flops / (kernel_ms * 1e-3) / 1e12, 3), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(flops / (kernel_ms * 1e-3) / PEAK_FP32_VALU, 4),
                     "traffic": None, "kernel_ms": round(kernel_ms, 3),
                     "attempted_steps": int(steps), "save_hit_steps": int(hits),
                     "note": "2,002 trajectories fill under 1 % of the lanes: latency-bound"},
    }
    if base is not None:
        out["cpu_baseline"] = base
        out["cpu_other_paths"] = others
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the ICs would shard as contiguous blocks, one per GPU, with one count reduction)",
              file=sys.stderr)
        sys.exit(2)
    workloads = {"surrogate": surrogate_workload, "duffing": duffing_workload,
                 "pendulum_c1": pendulum_c1_workload,
                 "lorenz": lorenz_workload, "estimator": estimator_workload,
                 "estimator_defaults": estimator_defaults_workload, "hdbscan": hdbscan_workload,
                 "mlp_infer": mlp_infer_workload, "rossler": rossler_workload,
                 "window_features": window_features_workload,
                 "window_spectrum": window_spectrum_workload,
                 "comprehensive": comprehensive_workload}
    workloads[args.workload](args)


if __name__ == "__main__":
    main()
