"""pybasin_amd: MI355X-native basin-stability integration path.

A ``HipDopri5Solver`` that plugs into pybasin's solver interface, HIP device
functors for the case-study systems, fused steady-state features, the
Rossler-network MFMA coupling kernel and a learned surrogate classifier.
"""

import logging

from .ode_system import ODESystem
from .solvers import FusedResult, HipDopri5Solver, Solver, TorchDopri5Solver

logging.getLogger("pybasin_amd").addHandler(logging.NullHandler())

__all__ = ["FusedResult", "HipDopri5Solver", "ODESystem", "Solver", "TorchDopri5Solver"]
__version__ = "0.1.0"
