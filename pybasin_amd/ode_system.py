"""ODESystem: the reference's torch ODE interface, with an optional device functor.

Mirrors ``ODESystem`` of src/pybasin/solvers/torch_ode_system.py:12-79:
an ``nn.Module`` whose ``ode(t, y, p)`` is the right-hand side, whose
``params_to_array()`` flattens ``self.params`` in declaration order, and
whose ``forward(t, y)`` reads per-row ``_batched_params`` when a solver sets
them.  A system that also names a device functor (``device_system_id``) is
integrated by the HIP kernel; any other system runs on the torch Dopri5.
"""

from __future__ import annotations

import ast
import inspect
import textwrap
from abc import ABC, abstractmethod
from typing import Any

import torch
import torch.nn as nn


class ODESystem(nn.Module, ABC):
    #: id of the HIP device functor (include/basinhip.h BH_SYS_*), or None
    device_system_id: int | None = None

    def __init__(self, params: dict[str, float]) -> None:
        super().__init__()
        self.params = params

    @abstractmethod
    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        """Right-hand side; p[..., i] is the i-th parameter (batched or flat)."""

    def params_to_array(self) -> torch.Tensor:
        return torch.tensor(list(dict(self.params).values()))

    def forward(self, t: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        p = getattr(self, "_batched_params", None)
        if p is None:
            p = self.params_to_array().to(y.device)
        return self.ode(t, y, p)

    def get_str(self) -> str:
        """Source of ``ode`` normalised through the AST plus the parameters
        (the cache-key string of src/pybasin/utils.py:56-89)."""
        try:
            src = textwrap.dedent(inspect.getsource(type(self).ode))
            body = ast.dump(ast.parse(src))
        except (OSError, TypeError):
            body = type(self).__name__
        return f"{type(self).__name__}:{body}:{dict(self.params)}"


def ode_param_count(ode: ODESystem) -> int:
    return len(dict(ode.params))


def describe(ode: Any) -> str:
    return getattr(ode, "get_str", lambda: type(ode).__name__)()
