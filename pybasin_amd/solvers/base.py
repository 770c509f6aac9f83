"""Solver base class with the reference's integrate() contract.

Restates ``Solver`` of src/pybasin/solvers/base.py:16-213:
  * t_eval must lie inside t_span (ValueError, base.py:54-55);
  * device auto-detection, "cuda" normalised to "cuda:0" (base.py:70-92);
  * save grid = torch.linspace(save_start, save_end, t_steps) in y0's dtype
    on the solver device (base.py:113-127);
  * y0 must be 2-D (ValueError, base.py:146-150);
  * params (P, n_p) expand IC-major: trajectory ic*P + p carries
    (y0[ic], params[p]) (base.py:155-159);
  * results optionally cached on disk (cache_manager.py:27-128).
"""

from __future__ import annotations

import logging
from abc import ABC, abstractmethod
from typing import Any

import torch

from ..cache import CacheManager

logger = logging.getLogger("pybasin_amd")

DEFAULT_CACHE_DIR = ".pybasin_cache"
UNSET = object()


class Solver(ABC):
    display_name = "Solver"

    def __init__(
        self,
        t_span: tuple[float, float] = (0, 1000),
        t_steps: int = 1000,
        device: str | None = None,
        rtol: float = 1e-8,
        atol: float = 1e-6,
        cache_dir: str | None = DEFAULT_CACHE_DIR,
        t_eval: tuple[float, float] | None = None,
    ):
        if t_eval is not None and (t_eval[0] < t_span[0] or t_eval[1] > t_span[1]):
            raise ValueError(f"t_eval {t_eval} must be within t_span {t_span}.")
        self.t_span = t_span
        self.t_eval = t_eval
        self.rtol = rtol
        self.atol = atol
        self.t_steps = t_steps
        self._set_device(device)
        self._cache_dir = cache_dir
        self._cache_manager: CacheManager | None = CacheManager(cache_dir) if cache_dir else None

    def _set_device(self, device: str | None) -> None:
        self._device_str = device
        if device is None:
            self.device = torch.device("cuda:0" if torch.cuda.is_available() else "cpu")
        else:
            dev = torch.device(device)
            if dev.type == "cuda":
                self.device = torch.device(f"cuda:{dev.index if dev.index is not None else 0}")
            else:
                self.device = dev

    @abstractmethod
    def clone(self, *, device: str | None = None, t_steps_factor: int = 1,
              cache_dir: str | None | object = UNSET) -> "Solver":
        ...

    def _prepare_tensors(self, y0: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        save_start, save_end = self.t_eval if self.t_eval is not None else self.t_span
        save_ts = torch.linspace(save_start, save_end, self.t_steps, dtype=y0.dtype,
                                 device=self.device)
        if y0.device != self.device:
            logger.warning("y0 is on device %s but solver expects %s", y0.device, self.device)
        return save_ts, y0

    def _expand(self, y0: torch.Tensor, params: torch.Tensor | None):
        if y0.ndim != 2:
            raise ValueError(
                f"y0 must be 2D with shape (batch, n_dims), got shape {tuple(y0.shape)}. "
                "For single initial condition, use y0.unsqueeze(0) or y0.reshape(1, -1).")
        if params is not None:
            if params.ndim != 2:
                raise ValueError(f"params must be 2-D (P, n_params), got {tuple(params.shape)}")
            B, P = y0.shape[0], params.shape[0]
            y0 = torch.repeat_interleave(y0, P, dim=0)
            params = params.to(self.device).repeat(B, 1)
        return y0, params

    def integrate(self, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None
                  ) -> tuple[torch.Tensor, torch.Tensor]:
        """Return (save_ts, y) with y of shape (t_steps, B*P, S)."""
        y0, params = self._expand(y0, params)
        save_ts, y0 = self._prepare_tensors(y0)
        ode_system = ode_system.to(self.device)

        def compute():
            return self._integrate(ode_system, y0, save_ts, params)

        if self._cache_manager is not None:
            return self._cache_manager.cached_call(
                solver_name=type(self).__name__, ode_system=ode_system, y0=y0, save_ts=save_ts,
                solver_config=self._get_cache_config(), device=self.device, compute_fn=compute,
                params=params)
        return compute()

    def _get_cache_config(self) -> dict[str, Any]:
        return {"rtol": self.rtol, "atol": self.atol, "t_span": self.t_span}

    @abstractmethod
    def _integrate(self, ode_system, y0: torch.Tensor, save_ts: torch.Tensor,
                   params: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        ...
