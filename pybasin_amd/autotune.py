"""Launch-shape autotuner for the integration kernels.

The tunables are the resident-lane cap of the per-lane kernel (how many
waves per SIMD stay resident; finished lanes refill from the counter) and
the resident-workgroup cap of the Rossler tile kernel.  The tuner times each
candidate on a short-horizon probe of the caller's own problem (same system,
dtype, parameters and a sample of its ICs), keeps the fastest, and caches
the choice per (device, system, dtype, size bucket) in a JSON file so later
runs reuse it without timing.
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


def tune_lane_cap(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                  probe_span: float | None = None, candidates=(0, 3, 2, 1),
                  use_cache: bool = True) -> int:
    """Best resident-lane cap (bh_solve_batch max_lanes; 0 = occupancy limit)
    for this problem.  candidates other than 0 are waves per SIMD."""
    from .solvers.hip_dopri5 import functor_id

    sid = functor_id(ode_system)
    if sid is None:
        raise ValueError("no device functor for this ODE system")
    dev = solver.device
    n = y0.shape[0] * (1 if params is None else params.shape[0])
    key = _key(dev, sid, y0.dtype, n)
    with _lock:
        cache = _load() if use_cache else {}
        if key in cache:
            return int(cache[key]["max_lanes"])
    props = torch.cuda.get_device_properties(dev)
    wave_set = props.multi_processor_count * 4 * 64
    span = probe_span if probe_span is not None else (solver.t_span[1] - solver.t_span[0]) / 20
    results = {}
    for c in candidates:
        lanes = 0 if c == 0 else int(c * wave_set)
        probe = type(solver)(t_span=(solver.t_span[0], solver.t_span[0] + span), t_steps=2,
                             device=str(dev), rtol=solver.rtol, atol=solver.atol, cache_dir=None,
                             stop_threshold=solver.stop_threshold, inf_fill=solver.inf_fill,
                             max_lanes=lanes)
        results[lanes] = _time(lambda: probe.integrate_fused(ode_system, y0, params))
    best = min(results, key=results.get)
    last_timings.clear()
    last_timings.update({str(k): round(v, 3) for k, v in results.items()})
    with _lock:
        cache = _load() if use_cache else {}
        cache[key] = {"max_lanes": best, "timings_ms": {str(k): v for k, v in results.items()}}
        if use_cache:
            _store(cache)
    return best


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
