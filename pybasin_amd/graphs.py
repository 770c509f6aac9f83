"""HIP-graph capture of a fused per-lane integration for repeated launches.

A basin-stability sweep launches the same integration shape many times (new
ICs, same sizes).  ``CapturedIntegration`` does the host work of a launch once
(validation, parameter and save-grid uploads), records the refill-counter
reset and the integration kernel (both phases and the compaction counter
of a two-phase launch) into a graph through torch's graph API, and replays
it after copying new ICs into the captured input buffer.  The counter is
reset by a kernel, never by a memset, so every replay starts from the same
state; outputs are overwritten on every replay.
"""

from __future__ import annotations

import torch

from .solvers.hip_dopri5 import FusedResult


class CapturedIntegration:
    def __init__(self, solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                 time_steady: float | None = None, keep_trajectories: bool = False):
        dev = solver.device
        if dev.type != "cuda":
            raise RuntimeError("graph capture needs a HIP device")
        self.solver = solver
        y0e, pe = solver._expand(y0.to(dev), params)
        save_ts, y0e = solver._prepare_tensors(y0e)
        self.plan = solver.plan_fused(ode_system, y0e.contiguous().clone(), save_ts, pe,
                                      time_steady=time_steady)
        self.keep = keep_trajectories
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # warm-up outside the graph (library load, allocator)
            solver.launch_fused(self.plan, keep_trajectories=self.keep)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.result: FusedResult = solver.launch_fused(self.plan, keep_trajectories=self.keep)

    @property
    def y0(self) -> torch.Tensor:
        return self.plan.y0

    def replay(self, y0: torch.Tensor | None = None) -> FusedResult:
        """Integrate again, optionally from new ICs of the captured shape
        (IC-major rows, already expanded when parameters are swept)."""
        if y0 is not None:
            if y0.shape != self.plan.y0.shape:
                raise ValueError(f"captured shape {tuple(self.plan.y0.shape)}, got {tuple(y0.shape)}")
            self.plan.y0.copy_(y0)
        self.graph.replay()
        return self.result
