"""Linear test systems (the reference's solver unit-test ODE and a 2x2 linear flow)."""

from __future__ import annotations

import torch

from .._lib import BH_SYS_DECAY, BH_SYS_LINEAR2
from ..ode_system import ODESystem


class ExponentialDecayODE(ODESystem):
    """dy/dt = decay * y (tests/unit/solvers/test_solver.py:12-21); params {"decay"}."""

    device_system_id = BH_SYS_DECAY

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        return p[..., 0:1] * y


class Linear2ODE(ODESystem):
    """y' = [[m00, m01], [m10, m11]] y; params {"m00", "m01", "m10", "m11"}."""

    device_system_id = BH_SYS_LINEAR2

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        d0 = p[..., 0] * y[..., 0] + p[..., 1] * y[..., 1]
        d1 = p[..., 2] * y[..., 0] + p[..., 3] * y[..., 1]
        return torch.stack([d0, d1], dim=-1)
