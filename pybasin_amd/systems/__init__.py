"""Case-study right-hand sides with HIP device functors.

Each class keeps the reference's torch ``ode`` (used by the torch Dopri5 and
by CPU clones) and names the matching device functor id from
include/basinhip.h.
"""

from .decay import ExponentialDecayODE, Linear2ODE
from .duffing import DuffingODE
from .lorenz import LorenzODE
from .pendulum import PendulumODE
from .rossler_network import RosslerNetworkODE, rossler_edges, rossler_laplacian

__all__ = [
    "DuffingODE", "ExponentialDecayODE", "Linear2ODE", "LorenzODE", "PendulumODE",
    "RosslerNetworkODE", "rossler_edges", "rossler_laplacian",
]
