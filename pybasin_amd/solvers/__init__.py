from .base import Solver
from .hip_dopri5 import (FUSED_RESULTS, FusedResult, HipDopri5Solver, functor_id,
                         register_functor, steady_index)
from .torch_dopri5 import TorchDopri5Solver, dopri5_shared

__all__ = ["FUSED_RESULTS", "FusedResult", "HipDopri5Solver", "Solver", "TorchDopri5Solver",
           "dopri5_shared", "functor_id", "register_functor", "steady_index"]
