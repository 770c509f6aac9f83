"""Launch-shape autotuner for the integration kernels.

The tunables are the resident-lane cap of the per-lane kernel (how many
waves per SIMD stay resident; finished lanes refill from the counter), its
refill quorum (how many finished lanes of a wave wait before they refill
together), the alternative two-phase launch with its step budget (long
trajectories packed into dense waves after the budget), and the
resident-workgroup cap of the Rossler tile kernel.  The per-lane tuner is
the library's bh_autotune_lane, which times each candidate (HIP events on
the launch stream) on a probe of the caller's own problem (same system, dtype,
parameters and ICs; by default the whole run with its save grid, since
early exits such as divergence stops decide whether refilling pays and may
come late; no trajectory output), keeps the fastest, and caches the choice per (device,
system, dtype, size bucket) in a JSON file so later runs reuse it without
timing.
"""

from __future__ import annotations

import json
import math
import os
import threading

import torch

from . import _lib

_CACHE_ENV = "PYBASIN_AMD_TUNE_CACHE"
_lock = threading.Lock()
last_timings: dict = {}  # candidate -> ms of the most recent tuning run


def _cache_path() -> str:
    return os.environ.get(_CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache",
                                                    "pybasin_amd_tune.json"))


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    path = _cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _key(device, system_id: int, dtype: torch.dtype, n: int) -> str:
    name = torch.cuda.get_device_name(device)
    return f"{name}|sys{system_id}|{dtype}|n2^{int(math.log2(max(n, 1)))}"


def _time(fn, reps: int = 2) -> float:
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_launch_shape(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                      probe_span: float | None = None, caps=(0, 3, 2, 1), quorums=(64, 1),
                      phase_quantiles=(0.25, 0.35, 0.45, 0.55, 0.65, 0.75, 0.8, 0.85),
                      use_cache: bool = True) -> dict:
    """Best launch shape for this problem: {"max_lanes", "refill_quorum",
    "phase_iters"} (set them on the solver; phase_iters > 0 selects the
    two-phase launch, else the refilling launch with max_lanes and
    refill_quorum).  Refill candidates: caps (waves per SIMD, 0 = occupancy
    limit) x quorums.  Phase candidates: the step budgets at which the given
    fractions of the trajectories have ended, read from the attempted-step
    counts of the first solve.  Without probe_span each candidate runs the
    solver's own span and save grid (without trajectory output); with it, a
    shorter run over [t0, t0 + probe_span] with two saves."""
    from .solvers.hip_dopri5 import functor_id

    sid = functor_id(ode_system)
    if sid is None:
        raise ValueError("no device functor for this ODE system")
    dev = solver.device
    n = y0.shape[0] * (1 if params is None else params.shape[0])
    key = (_key(dev, sid, y0.dtype, n) + f"|q{','.join(str(int(q)) for q in quorums)}"
           + f"|c{','.join(str(c) for c in caps)}|p{','.join(str(q) for q in phase_quantiles)}")
    with _lock:
        cache = _load() if use_cache else {}
        if key in cache:
            hit = cache[key]
            return {k: int(hit[k]) for k in ("max_lanes", "refill_quorum", "phase_iters")}
    props = torch.cuda.get_device_properties(dev)
    wave_set = props.multi_processor_count * 4 * 64
    lanes = [0 if c == 0 else int(c * wave_set) for c in caps]
    if probe_span is None:
        # the caller's own run: same span and save grid, no trajectories
        probe = solver.clone(cache_dir=None)
    else:
        probe = type(solver)(t_span=(solver.t_span[0], solver.t_span[0] + probe_span), t_steps=2,
                             device=str(dev), rtol=solver.rtol, atol=solver.atol, cache_dir=None,
                             stop_threshold=solver.stop_threshold, inf_fill=solver.inf_fill)
    y0e, pe = probe._expand(y0, params)
    save_ts, y0e = probe._prepare_tensors(y0e)
    plan = probe.plan_fused(ode_system, y0e, save_ts, pe)
    # bh_autotune_lane: one untimed and two timed launches per candidate
    # (quorum 0 is the library default, 64)
    _, _, _, results, res = probe.tune_fused(plan, lanes, [int(q) or 64 for q in quorums], reps=2)
    if phase_quantiles and res.nsteps.numel():
        steps = res.nsteps.double()
        top = int(steps.max())
        # torch.quantile takes at most 2^24 values: a strided sample of large batches
        sample = steps[:: max(1, steps.numel() // 1_000_000)].contiguous()
        qs = torch.quantile(sample, torch.tensor(list(phase_quantiles), dtype=torch.float64,
                                                 device=steps.device))
        budgets = sorted({int(v) + 1 for v in qs.cpu().tolist() if 0 < int(v) + 1 < top})
        if budgets:
            results.update(probe.tune_fused(plan, [], [], budgets, reps=2)[3])
    best = min(results, key=results.get)
    shape = ({"max_lanes": 0, "refill_quorum": 0, "phase_iters": best[1]} if best[0] == "phased"
             else {"max_lanes": best[1], "refill_quorum": best[2], "phase_iters": 0})
    names = {k: (f"phase{k[1]}" if k[0] == "phased" else f"{k[1]}/q{k[2]}") for k in results}
    last_timings.clear()
    last_timings.update({names[k]: round(v, 3) for k, v in results.items()})
    with _lock:
        cache = _load() if use_cache else {}
        cache[key] = dict(shape, timings_ms={names[k]: v for k, v in results.items()})
        if use_cache:
            _store(cache)
    return shape


def apply_launch_shape(solver, shape: dict) -> None:
    """Set a tuned launch shape (tune_launch_shape) on a HipDopri5Solver."""
    solver.max_lanes = int(shape["max_lanes"])
    solver.refill_quorum = int(shape["refill_quorum"])
    solver.phase_iters = int(shape["phase_iters"])


def tune_lane_launch(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                     probe_span: float | None = None, caps=(0, 3, 2, 1), quorums=(64, 1),
                     use_cache: bool = True) -> tuple[int, int]:
    """Best (resident-lane cap, refill quorum) pair of the refilling launch
    (tune_launch_shape without phase candidates)."""
    shape = tune_launch_shape(solver, ode_system, y0, params, probe_span, caps, quorums, (),
                              use_cache)
    return shape["max_lanes"], shape["refill_quorum"]


def tune_lane_cap(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                  probe_span: float | None = None, candidates=(0, 3, 2, 1),
                  use_cache: bool = True) -> int:
    """Best resident-lane cap at the solver's own refill quorum (see
    tune_lane_launch)."""
    return tune_lane_launch(solver, ode_system, y0, params, probe_span, caps=candidates,
                            quorums=(solver.refill_quorum,), use_cache=use_cache)[0]


def tune_rossler_tiles(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                       probe_span: float = 5.0, use_cache: bool = True) -> int:
    """Best resident-workgroup cap for the Rossler tile kernel (16 trajectories
    per workgroup; 0 = one per CU at the kernel's occupancy)."""
    dev = solver.device
    n = y0.shape[0] * (1 if params is None else params.shape[0])
    key = _key(dev, _lib.BH_SYS_ROSSLER100, y0.dtype, n)
    with _lock:
        cache = _load() if use_cache else {}
        if key in cache:
            return int(cache[key]["max_tiles"])
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    results = {}
    for tiles in (0, cus // 2, cus // 4):
        probe = type(solver)(t_span=(solver.t_span[0], solver.t_span[0] + probe_span), t_steps=2,
                             device=str(dev), rtol=solver.rtol, atol=solver.atol, cache_dir=None,
                             stop_threshold=solver.stop_threshold, max_lanes=16 * tiles)
        results[tiles] = _time(lambda: probe.integrate_fused(ode_system, y0, params))
    best = min(results, key=results.get)
    with _lock:
        cache = _load() if use_cache else {}
        cache[key] = {"max_tiles": best, "timings_ms": {str(k): v for k, v in results.items()}}
        if use_cache:
            _store(cache)
    return best
