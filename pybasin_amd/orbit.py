"""Orbit data: peak amplitudes per state over the steady window.

``extract_orbit_data`` follows src/pybasin/utils.py:254-303 on saved
trajectories (peaks = strict local maxima over a window of 2 n + 1 points,
edges excluded, torch_features_pattern.py:64-150; NaN-padded
(max_peaks, B, D) values in time order and (B, D) counts).

``orbit_data_from_result`` reads the same quantities from a fused launch
(bh_solve_batch_peaks, support 1), so no trajectory reaches HBM; and
``integrate_orbit`` runs that launch, re-running once with the exact
capacity when some trajectory has more peaks than the first buffer holds.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .features import DEFAULT_STEADY_FRACTION


@dataclass
class OrbitData:
    peak_values: torch.Tensor   # (max_peaks, B, D), NaN beyond each count
    peak_counts: torch.Tensor   # (B, D) int64
    dof_indices: list[int]
    time_steady: float


def find_peak_mask(x: torch.Tensor, n: int = 1) -> torch.Tensor:
    """(N, B, S) bool: x[k] > every value within n steps on either side;
    the first and last n points are never peaks."""
    N = x.shape[0]
    mask = torch.zeros(x.shape, dtype=torch.bool, device=x.device)
    if 2 * n >= N:
        return mask
    c = x[n:N - n]
    ok = torch.ones_like(c, dtype=torch.bool)
    for d in range(1, n + 1):
        ok &= (c > x[n - d:N - n - d]) & (c > x[n + d:N - n + d])
    mask[n:N - n] = ok
    return mask


def extract_peak_values(x: torch.Tensor, n: int = 1) -> tuple[torch.Tensor, torch.Tensor]:
    mask = find_peak_mask(x, n)
    counts = mask.sum(0)
    m = int(counts.max()) if counts.numel() else 0
    vals = torch.full((m,) + tuple(x.shape[1:]), float("nan"), dtype=x.dtype, device=x.device)
    if m:
        pos = mask.long().cumsum(0) - 1
        k, b, s = mask.nonzero(as_tuple=True)
        vals[pos[k, b, s], b, s] = x[k, b, s]
    return vals, counts


def _steady(t: torch.Tensor, time_steady: float | None) -> float:
    if time_steady is None:
        t0, t1 = float(t[0]), float(t[-1])
        return t0 + DEFAULT_STEADY_FRACTION * (t1 - t0)
    return float(time_steady)


def extract_orbit_data(t: torch.Tensor, y: torch.Tensor, dof: list[int],
                       time_steady: float | None = None, peak_support: int = 1) -> OrbitData:
    ts = _steady(t, time_steady)
    keep = t >= ts
    ys = y[keep][:, :, dof]
    if ys.shape[0] == 0:
        return OrbitData(torch.zeros((0, y.shape[1], len(dof)), dtype=y.dtype, device=y.device),
                         torch.zeros((y.shape[1], len(dof)), dtype=torch.long, device=y.device),
                         list(dof), ts)
    vals, counts = extract_peak_values(ys, peak_support)
    return OrbitData(vals, counts.long(), list(dof), ts)


def orbit_data_from_result(res, dof: list[int], time_steady: float | None = None) -> OrbitData:
    """OrbitData from a fused result launched with a peak buffer."""
    if res.peak_counts is None:
        raise ValueError("the fused result carries no peaks (launch with peak_cap)")
    counts = res.peak_counts[:, dof].long()
    m = int(counts.max()) if counts.numel() else 0
    if m > res.peak_values.shape[0]:
        raise ValueError(f"peak buffer holds {res.peak_values.shape[0]} peaks, a trajectory has {m}")
    vals = res.peak_values[:m][:, :, dof]
    k = torch.arange(m, device=vals.device)[:, None, None]
    vals = torch.where(k < counts[None], vals, torch.full_like(vals, float("nan")))
    return OrbitData(vals, counts, list(dof), _steady(res.t, time_steady))


def integrate_orbit(solver, ode_system, y0: torch.Tensor, params: torch.Tensor | None = None,
                    dof: list[int] | None = None, time_steady: float | None = None,
                    peak_cap: int = 64):
    """Fused integration with features and orbit peaks (support 1).  Returns
    (FusedResult, OrbitData)."""
    y0e, pe = solver._expand(y0.to(solver.device), params)
    save_ts, y0e = solver._prepare_tensors(y0e)
    plan = solver.plan_fused(ode_system, y0e, save_ts, pe, time_steady=time_steady)
    window = len(save_ts) - plan.steady_idx
    cap = max(0, min(int(peak_cap), (window - 1) // 2))
    res = solver.launch_fused(plan, keep_trajectories=False, peak_cap=cap)
    need = int(res.peak_counts.max()) if res.peak_counts.numel() else 0
    if need > cap:
        res = solver.launch_fused(plan, keep_trajectories=False, peak_cap=need)
    dof = list(range(y0e.shape[1])) if dof is None else list(dof)
    return res, orbit_data_from_result(res, dof, time_steady)
