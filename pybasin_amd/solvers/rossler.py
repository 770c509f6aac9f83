"""Host side of the Rossler-network tile kernel (bh_solve_rossler).

Only the state at the last save time is produced: full trajectories of the
configured sweep would be 110,000 x 1,000 x 300 float32 (132 GB), and the
synchronisation features the reference derives use the final step only
(case_studies/rossler_network/synchronization_feature_extractor.py:73-106).
"""

from __future__ import annotations

import torch

from .. import _lib
from ..systems.rossler_network import N_NODES, rossler_edge_arrays, rossler_laplacian


def run_rossler(solver, ode_system, y0: torch.Tensor, save_ts: torch.Tensor,
                params: torch.Tensor | None, *, keep_trajectories: bool,
                stream: torch.cuda.Stream | None = None, max_tiles: int = 0,
                coupling: str | None = None):
    """coupling: "edges" (the reference's edge-list form on the vector units,
    bh_solve_rossler_edges; the default) or "dense" (the Laplacian product on
    the matrix cores, bh_solve_rossler); the solver's rossler_coupling
    attribute sets the default."""
    from .hip_dopri5 import FusedResult, device_params

    if keep_trajectories:
        raise ValueError("the Rossler network kernel returns the final state only; "
                         "use integrate_fused(..., keep_trajectories=False)")
    if y0.dtype != torch.float32:
        raise ValueError("the Rossler network kernel runs in float32")
    dev = solver.device
    y0 = y0.to(dev).contiguous()
    n, S = y0.shape
    if S != 3 * N_NODES:
        raise ValueError(f"state dim must be {3 * N_NODES}, got {S}")
    p, p_stride = device_params(ode_system, params)
    p = p.to(dev).contiguous()
    if p_stride and p.shape[0] != n:
        raise ValueError(f"params rows {p.shape[0]} != trajectories {n}")
    coupling = coupling or getattr(solver, "rossler_coupling", "edges")
    if coupling not in ("edges", "dense"):
        raise ValueError(f"coupling must be 'edges' or 'dense', got {coupling!r}")
    final = torch.empty((n, S), dtype=torch.float32, device=dev)
    status = torch.empty((n,), dtype=torch.int32, device=dev)
    nsteps = torch.empty((n,), dtype=torch.int64, device=dev)
    ws = torch.empty((_lib.BH_WORKSPACE_BYTES,), dtype=torch.uint8, device=dev)
    flags, thr = 0, 0.0
    if solver.stop_threshold is not None:
        flags |= _lib.BH_FLAG_STOP
        thr = float(solver.stop_threshold)
        if solver.inf_fill:
            flags |= _lib.BH_FLAG_INF_FILL
    t_end = float(save_ts[-1])
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    tiles = int(max_tiles or getattr(solver, "max_lanes", 0) // 16)
    common = (float(solver.t_span[0]), t_end, float(solver.rtol), float(solver.atol),
              solver.max_steps, flags, thr, final.data_ptr(), status.data_ptr(), nsteps.data_ptr(),
              ws.data_ptr(), ws.numel(), tiles, s.cuda_stream)
    with torch.cuda.device(dev):
        if coupling == "edges":
            ei, ej = rossler_edge_arrays()  # host arrays, read by the call
            rc = _lib.lib().bh_solve_rossler_edges(
                _lib.BH_F32, y0.data_ptr(), n, p.data_ptr(), p_stride, ei.ctypes.data,
                ej.ctypes.data, len(ei), *common)
        else:
            lap = rossler_laplacian(torch.float32, dev).contiguous()
            rc = _lib.lib().bh_solve_rossler(
                _lib.BH_F32, y0.data_ptr(), n, p.data_ptr(), p_stride, lap.data_ptr(), *common)
    _lib.check(rc, "bh_solve_rossler_edges" if coupling == "edges" else "bh_solve_rossler")
    if stream is not None:  # the launch reads them on its own stream
        ws.record_stream(s)
        if coupling == "dense":
            lap.record_stream(s)
    return FusedResult(t=save_ts[-1:], y=None, final=final, stats=None, status=status,
                       nsteps=nsteps, steady_idx=len(save_ts) - 1)


def sync_features(final: torch.Tensor, n_nodes: int = N_NODES) -> torch.Tensor:
    """[max_dev_x, max_dev_y, max_dev_z, max_dev_all] of the final state
    (synchronization_feature_extractor.py:73-106)."""
    x, y, z = final[:, :n_nodes], final[:, n_nodes:2 * n_nodes], final[:, 2 * n_nodes:]
    dx = x.max(1).values - x.min(1).values
    dy = y.max(1).values - y.min(1).values
    dz = z.max(1).values - z.min(1).values
    return torch.stack([dx, dy, dz, torch.maximum(torch.maximum(dx, dy), dz)], dim=1)


def sync_labels(final: torch.Tensor, status: torch.Tensor, epsilon: float = 1.5):
    """'synchronized' if max deviation < epsilon (synchronization_classifier.py:36-54);
    trajectories stopped at the divergence threshold are 'unbounded'."""
    import numpy as np

    f = sync_features(final)[:, 3]
    lab = np.where((f < epsilon).cpu().numpy(), "synchronized", "desynchronized").astype(object)
    lab[(status == _lib.BH_LANE_DIVERGED).cpu().numpy()] = "unbounded"
    return lab
