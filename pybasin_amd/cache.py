"""On-disk cache of integration results (src/pybasin/cache_manager.py:27-128).

Key = md5 over the solver name, the ODE description, the ODE's parameter
values, the solver config, the bytes of y0 / save_ts / params; value = a safetensors file holding t and y.
A corrupt file is deleted and the result recomputed.

Fused results (``HipDopri5Solver.integrate_fused``) are cached the same way
under a key that also covers the fused options (feature window, kept
trajectories, orbit-peak capacity): the file holds the window statistics,
final states, status and step counts (and the trajectories / peaks when
kept), so a cached run does not touch the GPU.
"""

from __future__ import annotations

import hashlib
import json
import logging
import os

import torch

logger = logging.getLogger("pybasin_amd")


def _plain(v):
    """A parameter value as a hashable plain value (tensors by their values)."""
    if isinstance(v, torch.Tensor):
        return (str(v.dtype), tuple(v.detach().cpu().reshape(-1).tolist()))
    if isinstance(v, (list, tuple)):
        return tuple(_plain(x) for x in v)
    return v


class CacheManager:
    def __init__(self, cache_dir: str):
        self.cache_dir = cache_dir

    def build_key(self, solver_name, ode_system, y0, save_ts, solver_config, params=None) -> str:
        h = hashlib.md5()
        h.update(solver_name.encode())
        h.update(getattr(ode_system, "get_str", lambda: type(ode_system).__name__)().encode())
        # the parameter values, sorted by name (cache_manager.py:46-48): an
        # ODE's description may carry no parameters (utils.py:56-89)
        params_dict = dict(getattr(ode_system, "params", None) or {})
        h.update(repr(tuple(sorted((str(k), _plain(v)) for k, v in params_dict.items()))).encode())
        h.update(json.dumps(solver_config, sort_keys=True, default=str).encode())
        for t in (y0, save_ts, params):
            if t is not None:
                h.update(str(t.dtype).encode())
                h.update(t.detach().cpu().contiguous().numpy().tobytes())
        return h.hexdigest()

    def _path(self, key: str) -> str:
        return os.path.join(self.cache_dir, f"{key}.safetensors")

    def load(self, key: str, device):
        from safetensors.torch import load_file

        path = self._path(key)
        if not os.path.exists(path):
            return None
        try:
            d = load_file(path)
            return d["t"].to(device), d["y"].to(device)
        except Exception as exc:  # corrupt cache entry: drop it and recompute
            logger.warning("removing corrupt cache file %s (%s)", path, exc)
            os.remove(path)
            return None

    def save(self, key: str, t: torch.Tensor, y: torch.Tensor) -> None:
        from safetensors.torch import save_file

        os.makedirs(self.cache_dir, exist_ok=True)
        save_file({"t": t.detach().cpu().contiguous(), "y": y.detach().cpu().contiguous()},
                  self._path(key))

    # ---- fused results ------------------------------------------------------
    _FUSED_FIELDS = ("t", "y", "final", "stats", "status", "nsteps", "peak_values", "peak_counts")

    def save_fused(self, key: str, res) -> None:
        from safetensors.torch import save_file

        os.makedirs(self.cache_dir, exist_ok=True)
        tensors = {f: getattr(res, f).detach().cpu().contiguous() for f in self._FUSED_FIELDS
                   if getattr(res, f) is not None}
        save_file(tensors, self._path(key), metadata={"kind": "fused",
                                                      "steady_idx": str(int(res.steady_idx))})

    def load_fused(self, key: str, device):
        from safetensors import safe_open

        from .solvers.hip_dopri5 import FusedResult

        path = self._path(key)
        if not os.path.exists(path):
            return None
        try:
            with safe_open(path, framework="pt", device="cpu") as f:
                meta = f.metadata() or {}
                if meta.get("kind") != "fused":
                    raise ValueError("not a fused-result file")
                d = {k: f.get_tensor(k).to(device) for k in f.keys()}
            return FusedResult(t=d["t"], y=d.get("y"), final=d["final"], stats=d.get("stats"),
                               status=d["status"], nsteps=d["nsteps"],
                               steady_idx=int(meta["steady_idx"]),
                               peak_values=d.get("peak_values"), peak_counts=d.get("peak_counts"))
        except Exception as exc:
            logger.warning("removing corrupt cache file %s (%s)", path, exc)
            os.remove(path)
            return None

    def cached_call(self, solver_name, ode_system, y0, save_ts, solver_config, device,
                    compute_fn, params=None):
        key = self.build_key(solver_name, ode_system, y0, save_ts, solver_config, params)
        hit = self.load(key, device)
        if hit is not None:
            return hit
        t, y = compute_fn()
        self.save(key, t, y)
        return t, y
