"""Network of diffusively coupled Rossler oscillators.

Torch restatement of RosslerNetworkJaxODE
(case_studies/rossler_network/rossler_network_jax_ode.py:24-79) on the
100-node Watts-Strogatz graph of rossler_network_topology.py:11-14 (the edge
list is kept as data in ``data/rossler_ws100_edges.txt``).  State layout
[x_0..x_{N-1}, y_0..y_{N-1}, z_0..z_{N-1}]; per node
    x_i' = -y_i - z_i - K (L x)_i,   y_i' = x_i + a y_i,   z_i' = b + z_i (x_i - c)
with L = D - A the graph Laplacian (rossler_network_topology.py:64-77).
"""

from __future__ import annotations

import os
from functools import lru_cache

import numpy as np
import torch

from .._lib import BH_SYS_ROSSLER100
from ..ode_system import ODESystem

N_NODES = 100
N_EDGES = 400
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "rossler_ws100_edges.txt")


@lru_cache(maxsize=1)
def rossler_edges() -> np.ndarray:
    """(400, 2) int32 undirected edge list in file order."""
    rows = []
    with open(_DATA) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                i, j = line.split("-")
                rows.append((int(i), int(j)))
    e = np.asarray(rows, dtype=np.int32)
    assert e.shape == (N_EDGES, 2)
    return e


@lru_cache(maxsize=1)
def rossler_edge_arrays() -> tuple[np.ndarray, np.ndarray]:
    """(edges_i, edges_j), int32, both directions of every edge in file order:
    the arrays of the reference's build_edge_arrays
    (rossler_network_topology.py:26-48), (L x)_i = sum over e with
    edges_i[e] = i of (x_i - x_edges_j[e])."""
    e = rossler_edges()
    ei = np.ascontiguousarray(np.stack([e[:, 0], e[:, 1]], 1).reshape(-1), dtype=np.int32)
    ej = np.ascontiguousarray(np.stack([e[:, 1], e[:, 0]], 1).reshape(-1), dtype=np.int32)
    return ei, ej


@lru_cache(maxsize=1)
def _laplacian_np() -> np.ndarray:
    adj = np.zeros((N_NODES, N_NODES), dtype=np.float32)
    for i, j in rossler_edges():
        adj[i, j] = 1.0
        adj[j, i] = 1.0
    return np.diag(adj.sum(axis=1)) - adj


def rossler_laplacian(dtype=torch.float32, device=None) -> torch.Tensor:
    return torch.tensor(_laplacian_np(), dtype=dtype, device=device)


class RosslerNetworkODE(ODESystem):
    """params {"a", "b", "c", "K"}; the graph is the fixed 100-node network."""

    device_system_id = BH_SYS_ROSSLER100
    n_nodes = N_NODES

    def ode(self, t: torch.Tensor, y: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
        n = N_NODES
        a, b, c, k = p[..., 0:1], p[..., 1:2], p[..., 2:3], p[..., 3:4]
        x, ys, z = y[..., :n], y[..., n:2 * n], y[..., 2 * n:]
        lap = rossler_laplacian(y.dtype, y.device)
        coupling = x @ lap.T
        dx = -ys - z - k * coupling
        dy = x + a * ys
        dz = b + z * (x - c)
        return torch.cat([dx, dy, dz], dim=-1)
