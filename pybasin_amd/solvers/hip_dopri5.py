"""HipDopri5Solver: per-lane adaptive Dormand-Prince 5(4) on MI355X.

Drop-in ``Solver`` (src/pybasin/solvers/base.py:16-213, protocol in
src/pybasin/protocols.py:117-187) whose ``_integrate`` replaces
TorchDiffEqSolver._integrate (src/pybasin/solvers/torchdiffeq_solver.py:97-134).

* Systems with a device functor (``ode_system.device_system_id``) run in the
  HIP kernel of libbasinhip.so: one trajectory per lane, torchdiffeq's step
  rules per lane, streaming features fused into the save loop.  If the
  library is missing this raises; there is no silent fallback.
* The 100-node Rossler network runs in the tile kernel (MFMA coupling GEMM).
* Systems without a functor, and solvers cloned to the CPU
  (TemplateIntegrator, src/pybasin/template_integrator.py:79-82), run on the
  torch shared-step Dopri5 (``TorchDopri5Solver``) -- by construction, logged.

The solver holds no per-call state, so two threads may call ``integrate``
concurrently (the reference's main + template integration,
src/pybasin/basin_stability_estimator.py:498-512); ctypes releases the GIL
during the launch.
"""

from __future__ import annotations

import logging
import threading
from dataclasses import dataclass
from typing import Any

import torch
from torch.utils.weak import WeakTensorKeyDictionary

from .. import _lib
from .base import DEFAULT_CACHE_DIR, UNSET, Solver
from .torch_dopri5 import TorchDopri5Solver

logger = logging.getLogger("pybasin_amd")

# fused results attached to returned trajectory tensors (read by
# HipFusedFeatureExtractor); keyed weakly so nothing outlives the tensor, and
# by tensor identity: a plain WeakKeyDictionary compares keys with
# Tensor.__eq__, which is elementwise and cannot be a dictionary test
FUSED_RESULTS: "WeakTensorKeyDictionary" = WeakTensorKeyDictionary()


# ODE classes that do not carry ``device_system_id`` themselves (e.g. the
# reference's own case-study classes) opt in here; registration is explicit
# so a class is never matched to a functor by name alone.
_FUNCTORS: "dict[type, int]" = {}


def register_functor(ode_class: type, system_id: int) -> None:
    """Integrate instances of ode_class with the HIP functor system_id
    (include/basinhip.h BH_SYS_*); the class's params must follow that
    functor's parameter order."""
    _FUNCTORS[ode_class] = int(system_id)


def functor_id(ode_system) -> int | None:
    sid = getattr(ode_system, "device_system_id", None)
    if sid is not None:
        return int(sid)
    for cls in type(ode_system).__mro__:
        if cls in _FUNCTORS:
            return _FUNCTORS[cls]
    return None


def device_params(ode_system, params: torch.Tensor | None) -> tuple[torch.Tensor, int]:
    """float64 parameter block for the kernels: (rows, n_p) with stride n_p,
    or the system's own parameters as one row with stride 0 (built from the
    params dict in declaration order, as params_to_array does)."""
    if params is None:
        vals = [float(v) for v in dict(ode_system.params).values()]
        return torch.tensor([vals], dtype=torch.float64), 0
    p = params.to(torch.float64).contiguous()
    return p, p.shape[1]


# Device copies of shared parameter vectors, by (device, values): a fresh
# host-to-device copy from pageable memory would synchronise the stream on
# every launch.  The cached tensors are only read by the kernels.
_PARAM_CACHE: dict = {}
_PARAM_CACHE_LOCK = threading.Lock()


def _shared_params_on(p: torch.Tensor, dev: torch.device) -> torch.Tensor:
    key = (str(dev), tuple(p.flatten().tolist()))
    with _PARAM_CACHE_LOCK:
        hit = _PARAM_CACHE.get(key)
        if hit is None:
            if len(_PARAM_CACHE) >= 256:
                _PARAM_CACHE.clear()
            hit = _PARAM_CACHE[key] = p.to(dev).contiguous()
    return hit


@dataclass
class LanePlan:
    """Device-resident inputs of one bh_solve_batch launch."""

    sys_id: int
    y0: torch.Tensor
    p: torch.Tensor
    p_stride: int
    save_ts: torch.Tensor
    st64: torch.Tensor
    steady_idx: int


@dataclass
class FusedResult:
    """Outputs of one fused HIP integration.

    t       (n_save,) save grid in the state dtype
    y       (n_save, n, S) trajectories, or None when not kept
    final   (n, S) state at the last save time
    stats   (n, S, 4) float64: mean, std (ddof 0), min, max over the window
            ((n, S, 5) with the least-squares slope when run with trend=True)
    peak_values / peak_counts  orbit peaks over the window when requested
    status  (n,) int32 lane status (BH_LANE_*)
    nsteps  (n,) int64 attempted steps per trajectory
    steady_idx  first save index of the feature window
    """

    t: torch.Tensor
    y: torch.Tensor | None
    final: torch.Tensor
    stats: torch.Tensor | None
    status: torch.Tensor
    nsteps: torch.Tensor
    steady_idx: int
    peak_values: torch.Tensor | None = None   # (peak_cap, n, S), see bh_solve_batch_peaks
    peak_counts: torch.Tensor | None = None   # (n, S) int32


def steady_index(save_ts: torch.Tensor, time_steady: float | None) -> int:
    """filter_time's start index (src/pybasin/feature_extractors/feature_extractor.py:85-130):
    first save time >= time_steady (default t0 + 0.85 (t1 - t0)); 0 if none."""
    if time_steady is None:
        t0, t1 = float(save_ts[0]), float(save_ts[-1])
        time_steady = t0 + 0.85 * (t1 - t0)
    idx = torch.nonzero(save_ts.cpu() >= time_steady)
    return int(idx[0, 0]) if len(idx) else 0


def count_save_steps(solver, ode_system, y0: torch.Tensor, save_ts: torch.Tensor,
                     params: torch.Tensor | None, time_steady: float | None = None, *,
                     steady_idx: int | None = None) -> tuple[int, int]:
    """Total attempted steps and total steps whose span held a save point
    (the dense-output fits) of solving this problem with the solver's launch
    shape: one extra launch with BH_FLAG_COUNT_SAVE_STEPS.  A trajectory's
    steps do not depend on the launch, so these are the counts of every
    launch of the same problem."""
    plan = solver.plan_fused(ode_system, y0, save_ts, params, time_steady=time_steady,
                             steady_idx=steady_idx)
    res = solver.launch_fused(plan, keep_trajectories=False,
                              extra_flags=_lib.BH_FLAG_COUNT_SAVE_STEPS)
    packed = res.nsteps
    steps = int((packed & 0xFFFFFFFF).sum())
    hits = int((packed >> 32).sum())
    return steps, hits


class HipDopri5Solver(Solver):
    display_name = "HIP Dopri5 (per lane)"

    def __init__(
        self,
        t_span: tuple[float, float] = (0, 1000),
        t_steps: int = 1000,
        device: str | None = None,
        rtol: float = 1e-8,
        atol: float = 1e-6,
        cache_dir: str | None = DEFAULT_CACHE_DIR,
        t_eval: tuple[float, float] | None = None,
        *,
        max_steps: int = 16**5,
        stop_threshold: float | None = None,
        inf_fill: bool = False,
        max_lanes: int = 0,
        refill_quorum: int = 0,
        phase_iters: int = 0,
    ):
        super().__init__(t_span, t_steps, device, rtol, atol, cache_dir, t_eval)
        self.max_steps = int(max_steps)
        self.stop_threshold = stop_threshold
        self.inf_fill = inf_fill
        self.max_lanes = int(max_lanes)
        if not 0 <= int(refill_quorum) <= 64:
            raise ValueError(f"refill_quorum must be in [0, 64], got {refill_quorum}")
        self.refill_quorum = int(refill_quorum)
        if int(phase_iters) < 0:
            raise ValueError(f"phase_iters must be >= 0, got {phase_iters}")
        self.phase_iters = int(phase_iters)

    def _get_cache_config(self) -> dict[str, Any]:
        cfg = super()._get_cache_config()
        cfg.update(method="hip_dopri5_lane", stop_threshold=self.stop_threshold,
                   inf_fill=self.inf_fill, max_steps=self.max_steps)
        return cfg

    def clone(self, *, device=None, t_steps_factor=1, cache_dir=UNSET) -> "HipDopri5Solver":
        return HipDopri5Solver(
            t_span=self.t_span, t_steps=self.t_steps * t_steps_factor,
            device=device or self._device_str, rtol=self.rtol, atol=self.atol,
            cache_dir=self._cache_dir if cache_dir is UNSET else cache_dir,  # type: ignore[arg-type]
            t_eval=self.t_eval, max_steps=self.max_steps, stop_threshold=self.stop_threshold,
            inf_fill=self.inf_fill, max_lanes=self.max_lanes, refill_quorum=self.refill_quorum,
            phase_iters=self.phase_iters)

    # ------------------------------------------------------------------ #
    def _torch_solver(self) -> TorchDopri5Solver:
        return TorchDopri5Solver(t_span=self.t_span, t_steps=self.t_steps,
                                 device=str(self.device), rtol=self.rtol, atol=self.atol,
                                 cache_dir=None, t_eval=self.t_eval)

    def _integrate(self, ode_system, y0, save_ts, params=None):
        sys_id = functor_id(ode_system)
        if self.device.type != "cuda" or sys_id is None:
            logger.info("[%s] %s on %s: torch shared-step Dopri5", type(self).__name__,
                        type(ode_system).__name__, self.device)
            return self._torch_solver()._integrate(ode_system, y0, save_ts, params)
        res = self.run_fused(ode_system, y0, save_ts, params, keep_trajectories=True)
        FUSED_RESULTS[res.y] = res
        return save_ts, res.y

    def integrate_fused(self, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                        *, time_steady: float | None = None,
                        keep_trajectories: bool = False, trend: bool = False) -> FusedResult:
        """integrate() with the feature window fused into the kernel; trajectories
        reach HBM only when keep_trajectories is set.  With a cache_dir the
        result is read from / written to the on-disk cache (a hit skips the
        GPU)."""
        if self.device.type != "cuda":
            raise RuntimeError("integrate_fused runs on a HIP device; use integrate() on CPU")
        y0, params = self._expand(y0, params)
        save_ts, y0 = self._prepare_tensors(y0)
        if self._cache_manager is None:
            return self.run_fused(ode_system, y0, save_ts, params,
                                  keep_trajectories=keep_trajectories, time_steady=time_steady,
                                  trend=trend)
        cfg = dict(self._get_cache_config(), fused=True,
                   steady_idx=steady_index(save_ts, time_steady),
                   keep_trajectories=bool(keep_trajectories), trend=bool(trend))
        key = self._cache_manager.build_key(type(self).__name__, ode_system, y0, save_ts, cfg,
                                            params)
        hit = self._cache_manager.load_fused(key, self.device)
        if hit is not None:
            return hit
        res = self.run_fused(ode_system, y0, save_ts, params, keep_trajectories=keep_trajectories,
                             time_steady=time_steady, trend=trend)
        self._cache_manager.save_fused(key, res)
        return res

    def run_fused(self, ode_system, y0: torch.Tensor, save_ts: torch.Tensor,
                  params: torch.Tensor | None, *, keep_trajectories: bool,
                  time_steady: float | None = None, steady_idx: int | None = None,
                  stream: torch.cuda.Stream | None = None, trend: bool = False) -> FusedResult:
        """One fused launch on prepared tensors.  trend=True adds the window's
        least-squares slope per state as a fifth statistic (BH_FLAG_TREND)."""
        sys_id = functor_id(ode_system)
        if sys_id is None:
            raise ValueError(f"{type(ode_system).__name__} has no HIP device functor "
                             "(set device_system_id or call register_functor)")
        if sys_id == _lib.BH_SYS_ROSSLER100:
            from .rossler import run_rossler  # MFMA tile kernel

            return run_rossler(self, ode_system, y0, save_ts, params,
                               keep_trajectories=keep_trajectories, stream=stream)
        plan = self.plan_fused(ode_system, y0, save_ts, params, time_steady=time_steady,
                               steady_idx=steady_idx)
        return self.launch_fused(plan, keep_trajectories=keep_trajectories, stream=stream,
                                 extra_flags=_lib.BH_FLAG_TREND if trend else 0)

    def plan_fused(self, ode_system, y0: torch.Tensor, save_ts: torch.Tensor,
                   params: torch.Tensor | None, *, time_steady: float | None = None,
                   steady_idx: int | None = None) -> "LanePlan":
        """Host-side preparation of a bh_solve_batch launch: validation and every
        host-to-device copy, so that launch_fused issues device work only (and
        can be captured into a graph)."""
        sys_id = functor_id(ode_system)
        if sys_id is None or sys_id == _lib.BH_SYS_ROSSLER100:
            raise ValueError(f"{type(ode_system).__name__} has no per-lane device functor")
        dev = self.device
        if y0.dtype not in (torch.float32, torch.float64):
            raise ValueError(f"y0 dtype must be float32 or float64, got {y0.dtype}")
        y0 = y0.to(dev).contiguous()
        n, S = y0.shape
        p, p_stride = device_params(ode_system, params)
        p = _shared_params_on(p, dev) if p_stride == 0 else p.to(dev).contiguous()
        if p_stride and p.shape[0] != n:
            raise ValueError(f"params rows {p.shape[0]} != trajectories {n}")
        st64 = save_ts.to(device=dev, dtype=torch.float64).contiguous()
        if steady_idx is None:
            steady_idx = steady_index(save_ts, time_steady)
        return LanePlan(sys_id=int(sys_id), y0=y0, p=p, p_stride=p_stride, save_ts=save_ts,
                        st64=st64, steady_idx=int(steady_idx))

    def _lane_call(self, plan: "LanePlan", keep_trajectories: bool, extra_flags: int = 0):
        """Output tensors and the bh_solve_batch arguments shared by every
        lane entry point (system id through out_nsteps), for a prepared plan."""
        dev = self.device
        y0, p, st64 = plan.y0, plan.p, plan.st64
        dtype_id = _lib.BH_F32 if y0.dtype == torch.float32 else _lib.BH_F64
        n, S = y0.shape
        n_save = st64.shape[0]
        nf = _lib.BH_NFEAT_TREND if extra_flags & _lib.BH_FLAG_TREND else _lib.BH_NFEAT
        out = dict(
            traj=torch.empty((n_save, n, S), dtype=y0.dtype, device=dev) if keep_trajectories else None,
            stats=torch.empty((n, S, nf), dtype=torch.float64, device=dev),
            final=torch.empty((n, S), dtype=y0.dtype, device=dev),
            status=torch.empty((n,), dtype=torch.int32, device=dev),
            nsteps=torch.empty((n,), dtype=torch.int64, device=dev),
            ws=torch.empty((_lib.BH_WORKSPACE_BYTES,), dtype=torch.uint8, device=dev))
        flags = ((self.refill_quorum & 0x7F) << _lib.BH_REFILL_QUORUM_SHIFT) | int(extra_flags)
        thr = 0.0
        if self.stop_threshold is not None:
            flags |= _lib.BH_FLAG_STOP
            thr = float(self.stop_threshold)
            if self.inf_fill:
                flags |= _lib.BH_FLAG_INF_FILL
        common = (plan.sys_id, dtype_id, y0.data_ptr(), n, S, p.data_ptr(), p.shape[1],
                  plan.p_stride, float(self.t_span[0]), st64.data_ptr(), n_save,
                  float(self.rtol), float(self.atol), self.max_steps, flags, thr, plan.steady_idx,
                  _lib.ptr(out["traj"]), out["stats"].data_ptr(), out["final"].data_ptr(),
                  out["status"].data_ptr(), out["nsteps"].data_ptr())
        return out, common

    def _result(self, plan: "LanePlan", out: dict, pk_vals=None, pk_counts=None) -> FusedResult:
        n_save = plan.st64.shape[0]
        return FusedResult(t=plan.save_ts, y=out["traj"], final=out["final"],
                           stats=out["stats"] if plan.steady_idx < n_save else None,
                           status=out["status"], nsteps=out["nsteps"],
                           steady_idx=int(plan.steady_idx), peak_values=pk_vals,
                           peak_counts=pk_counts)

    def launch_fused(self, plan: "LanePlan", *, keep_trajectories: bool,
                     stream: torch.cuda.Stream | None = None,
                     peak_cap: int | None = None, extra_flags: int = 0) -> FusedResult:
        """Allocate the outputs and launch bh_solve_batch for a prepared plan
        (bh_solve_batch_peaks with a peak buffer of peak_cap rows when
        peak_cap is given).  extra_flags are or-ed into the launch flags
        (BH_FLAG_COUNT_SAVE_STEPS)."""
        dev = self.device
        y0, p, st64 = plan.y0, plan.p, plan.st64
        n, S = y0.shape
        out, common = self._lane_call(plan, keep_trajectories, extra_flags)
        ws = out["ws"]
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        pk_vals = pk_counts = None
        with torch.cuda.device(dev):
            if peak_cap is None and self.phase_iters > 0:
                # two launches with the long trajectories packed in between
                need = _lib.lib().bh_solve_batch_phased_workspace_bytes(common[0], common[1], S, n)
                _lib.check(-5 if need < 0 else 0, "bh_solve_batch_phased_workspace_bytes")
                ws = out["ws"] = torch.empty((need,), dtype=torch.uint8, device=dev)
                rc = _lib.lib().bh_solve_batch_phased(*common, self.phase_iters, ws.data_ptr(),
                                                      ws.numel(), s.cuda_stream)
            elif peak_cap is None:
                rc = _lib.lib().bh_solve_batch(*common, ws.data_ptr(), ws.numel(),
                                               self.max_lanes, s.cuda_stream)
            else:
                cap = max(int(peak_cap), 0)
                pk_vals = torch.full((cap, n, S), float("nan"), dtype=y0.dtype, device=dev)
                pk_counts = torch.empty((n, S), dtype=torch.int32, device=dev)
                rc = _lib.lib().bh_solve_batch_peaks(*common, cap, _lib.ptr(pk_vals) if cap else None,
                                                     pk_counts.data_ptr(), ws.data_ptr(),
                                                     ws.numel(), self.max_lanes, s.cuda_stream)
        _lib.check(rc, "bh_solve_batch")
        for t in (y0, p, st64, ws):
            t.record_stream(s)
        return self._result(plan, out, pk_vals, pk_counts)

    def tune_fused(self, plan: "LanePlan", caps=(0,), quorums=(64,), phases=(), reps: int = 2,
                   stream: torch.cuda.Stream | None = None):
        """bh_autotune_lane on a prepared plan: solve it once per candidate --
        each (max_lanes cap, refill quorum) pair and each phase budget of the
        two-phase launch -- timing each launch, and return (best cap, best
        quorum, best phase_iters, {candidate: ms}, FusedResult of the problem);
        candidates are keyed ("refill", cap, quorum) and ("phased", K).  Blocks
        until the timed launches end; no trajectories are written."""
        import ctypes

        dev = self.device
        out, common = self._lane_call(plan, False)
        caps = [int(c) for c in caps]
        quorums = [int(q) for q in quorums]
        phases = [int(k) for k in phases]
        ws = out["ws"]
        if phases:
            n, S = plan.y0.shape
            need = _lib.lib().bh_solve_batch_phased_workspace_bytes(common[0], common[1], S, n)
            _lib.check(-5 if need < 0 else 0, "bh_solve_batch_phased_workspace_bytes")
            ws = out["ws"] = torch.empty((max(need, _lib.BH_WORKSPACE_BYTES),), dtype=torch.uint8,
                                         device=dev)
        arr = lambda v: (ctypes.c_int32 * max(len(v), 1))(*v)  # noqa: E731
        n_t = len(caps) * len(quorums) + len(phases)
        c_t = (ctypes.c_float * max(n_t, 1))()
        best_cap, best_q, best_k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        with torch.cuda.device(dev):
            rc = _lib.lib().bh_autotune_lane(*common, ws.data_ptr(), ws.numel(), arr(caps), len(caps),
                                             arr(quorums), len(quorums), arr(phases), len(phases),
                                             int(reps), ctypes.byref(best_cap), ctypes.byref(best_q),
                                             ctypes.byref(best_k), c_t, s.cuda_stream)
        _lib.check(rc, "bh_autotune_lane")
        timings = {("refill", c, q): float(c_t[i * len(quorums) + j])
                   for i, c in enumerate(caps) for j, q in enumerate(quorums)}
        base = len(caps) * len(quorums)
        timings.update({("phased", k): float(c_t[base + i]) for i, k in enumerate(phases)})
        return best_cap.value, best_q.value, best_k.value, timings, self._result(plan, out)
