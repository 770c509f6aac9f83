"""Duffing oscillator (case_studies/duffing_oscillator/duffing_ode.py:8-42)."""

from __future__ import annotations

import torch

from .._lib import BH_SYS_DUFFING
from ..ode_system import ODESystem


class DuffingODE(ODESystem):
    """x' = v; v' = -delta*v - k3*x^3 + A*cos(t); params {"delta", "k3", "A"}."""

    device_system_id = BH_SYS_DUFFING

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        delta, k3, amplitude = p[..., 0], p[..., 1], p[..., 2]
        x, x_dot = y[..., 0], y[..., 1]
        return torch.stack([x_dot, -delta * x_dot - k3 * x**3 + amplitude * torch.cos(t)], dim=-1)
