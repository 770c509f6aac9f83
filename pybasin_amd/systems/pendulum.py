"""Forced damped pendulum (case_studies/pendulum/pendulum_ode.py:10-40)."""

from __future__ import annotations

import torch

from .._lib import BH_SYS_PENDULUM
from ..ode_system import ODESystem


class PendulumODE(ODESystem):
    """theta' = omega; omega' = -alpha*omega + T - K*sin(theta); params {"alpha", "T", "K"}."""

    device_system_id = BH_SYS_PENDULUM

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        alpha, torque, k = p[..., 0], p[..., 1], p[..., 2]
        theta, theta_dot = y[..., 0], y[..., 1]
        return torch.stack([theta_dot, -alpha * theta_dot + torque - k * torch.sin(theta)], dim=-1)
