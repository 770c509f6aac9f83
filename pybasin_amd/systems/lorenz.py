"""Lorenz system with the freeze mask (case_studies/lorenz/lorenz_ode.py:8-64)."""

from __future__ import annotations

import torch

from .._lib import BH_SYS_LORENZ
from ..ode_system import ODESystem


class LorenzODE(ODESystem):
    """x' = sigma(y-x); y' = r x - x z - y; z' = x y - b z; the derivative is
    zeroed once max|state| >= 200; params {"sigma", "r", "b"}."""

    device_system_id = BH_SYS_LORENZ
    #: state magnitude at which the mask freezes a trajectory
    freeze_threshold = 200.0

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        sigma, r, b = p[..., 0], p[..., 1], p[..., 2]
        x, y_, z = y[..., 0], y[..., 1], y[..., 2]
        dydt = torch.stack([sigma * (y_ - x), r * x - x * z - y_, x * y_ - b * z], dim=-1)
        mask = (torch.max(torch.abs(y), dim=-1)[0] < 200).to(y.dtype).unsqueeze(-1)
        return dydt * mask
