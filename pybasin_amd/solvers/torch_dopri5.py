"""Torch Dormand-Prince 5(4) with one shared step size (torchdiffeq semantics).

The solver for ODE systems without a HIP device functor and for CPU clones
(TemplateIntegrator clones the solver to the CPU,
src/pybasin/template_integrator.py:79-82).  It runs the algorithm of
torchdiffeq 0.2.5's ``dopri5`` that the reference's TorchDiffEqSolver calls
(src/pybasin/solvers/torchdiffeq_solver.py:97-134): the whole (B, S) batch is
one tensor with one dt, the error norm is the RMS over all elements, time-like
values are float64, stage times are cast to the state dtype, and save times
are reached by dense output without clamping the last step.
"""

from __future__ import annotations

import logging
from typing import Any

import torch

from .base import DEFAULT_CACHE_DIR, UNSET, Solver

logger = logging.getLogger("pybasin_amd")

_A = [
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
_C = [1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]
_E = [35 / 384 - 1951 / 21600, 0.0, 500 / 1113 - 22642 / 50085, 125 / 192 - 451 / 720,
      -2187 / 6784 - -12231 / 42400, 11 / 84 - 649 / 6300, -1.0 / 60.0]
_MID = [6025192743 / 30085553152 / 2, 0.0, 51252292925 / 65400821598 / 2,
        -2691868925 / 45128329728 / 2, 187940372067 / 1594534317056 / 2,
        -1776094331 / 19743644256 / 2, 11237099 / 235043384 / 2]


def _rms(x: torch.Tensor) -> torch.Tensor:
    return x.abs().pow(2).mean().sqrt()


def dopri5_shared(func, y0: torch.Tensor, t0: float, save_ts: torch.Tensor, rtol: float,
                  atol: float, max_steps: int = 2**31 - 1) -> torch.Tensor:
    """Integrate y' = func(t, y) for the batch y0 (B, S) from t0; returns
    (len(save_ts), B, S).  save_ts must be non-decreasing and >= t0."""
    dt_ = y0.dtype
    dev = y0.device
    tsv = [float(v) for v in save_ts.to(torch.float64).cpu()]
    f = lambda t, y: func(torch.tensor(t, dtype=dt_, device=dev), y)  # noqa: E731
    A = [torch.tensor(r, dtype=dt_, device=dev) for r in _A]
    E = torch.tensor(_E, dtype=dt_, device=dev)
    MID = torch.tensor(_MID, dtype=dt_, device=dev)
    y = y0.clone()
    f0 = f(t0, y)
    # initial step (Hairer, Norsett & Wanner II.4)
    scale = atol + y.abs() * rtol
    d0, d1 = _rms(y / scale), _rms(f0 / scale)
    h0 = torch.tensor(1e-6, dtype=dt_, device=dev) if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    h0 = h0.abs()
    f1 = f(t0 + float(h0), y + h0 * f0)
    d2 = (_rms((f1 - f0) / scale) / h0).abs()
    if d1 <= 1e-15 and d2 <= 1e-15:
        h1 = torch.clamp(h0 * 1e-3, min=1e-6)
    else:
        h1 = (0.01 / torch.maximum(d1, d2)) ** (1.0 / 5.0)
    dt = float(torch.minimum(100 * h0, h1.abs()))
    out = torch.empty((len(tsv),) + tuple(y0.shape), dtype=dt_, device=dev)
    coef = [y, torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y), torch.zeros_like(y)]
    ti0 = ti1 = t0
    steps = 0
    for si, ts in enumerate(tsv):
        while ts > ti1:
            if steps >= max_steps:
                raise RuntimeError(f"max_num_steps exceeded ({steps})")
            tA, tB = ti1, ti1 + dt
            if not tB > tA:
                raise RuntimeError(f"underflow in dt {dt}")
            if not torch.isfinite(y).all():
                raise RuntimeError("non-finite values in state")
            dtf = torch.tensor(dt, dtype=dt_, device=dev)
            tB_f = torch.tensor(tB, dtype=dt_, device=dev)
            t_prev = torch.nextafter(tB_f, tB_f - 1)
            ks = [f0]
            for st in range(6):
                co = A[st] * dtf
                acc = ks[0] * co[0]
                for j in range(1, st + 1):
                    acc = acc + ks[j] * co[j]
                ys = y + acc
                tst = t_prev if _C[st] == 1.0 else torch.tensor(tA, dtype=dt_, device=dev) + _C[st] * dtf
                ks.append(func(tst, ys))
            y1 = ys
            ec = E * dtf
            err = ks[0] * ec[0]
            for j in range(1, 7):
                err = err + ks[j] * ec[j]
            tol = atol + rtol * torch.maximum(y.abs(), y1.abs())
            ratio = float(_rms(err / tol))
            steps += 1
            if ratio <= 1.0:
                mc = MID * dtf
                acc = ks[0] * mc[0]
                for j in range(1, 7):
                    acc = acc + ks[j] * mc[j]
                ymid = y + acc
                fa, fb = ks[0], ks[6]
                a = 2 * dtf * (fb - fa) - 8 * (y1 + y) + 16 * ymid
                b = dtf * (5 * fa - 3 * fb) + 18 * y + 14 * y1 - 32 * ymid
                c = dtf * (fb - 4 * fa) - 11 * y - 5 * y1 + 16 * ymid
                coef = [y, dtf * fa, c, b, a]
                y, f0 = y1, fb
                ti0, ti1 = tA, tB
            if ratio == 0.0:
                dt = dt * 10.0
            else:
                dfac = 1.0 if ratio < 1.0 else 0.2
                dt = dt * min(10.0, max(0.9 / ratio ** 0.2, dfac))
        x = (ts - ti0) / (ti1 - ti0) if ti1 > ti0 else 0.0
        x = torch.tensor(x, dtype=dt_, device=dev)
        tot = coef[0] + x * coef[1]
        xp = x
        for cf in coef[2:]:
            xp = xp * x
            tot = tot + xp * cf
        out[si] = tot
    return out


class TorchDopri5Solver(Solver):
    """Shared-step Dopri5 in torch for any ODESystem (CPU or GPU)."""

    display_name = "Torch Dopri5 (shared step)"

    def __init__(self, t_span=(0, 1000), t_steps=1000, device=None, rtol=1e-8, atol=1e-6,
                 cache_dir: str | None = DEFAULT_CACHE_DIR, t_eval=None,
                 max_steps: int = 2**31 - 1):
        super().__init__(t_span, t_steps, device, rtol, atol, cache_dir, t_eval)
        self.max_steps = max_steps

    def _get_cache_config(self) -> dict[str, Any]:
        cfg = super()._get_cache_config()
        cfg["method"] = "dopri5_shared"
        return cfg

    def clone(self, *, device=None, t_steps_factor=1, cache_dir=UNSET) -> "TorchDopri5Solver":
        return TorchDopri5Solver(
            t_span=self.t_span, t_steps=self.t_steps * t_steps_factor,
            device=device or self._device_str, rtol=self.rtol, atol=self.atol,
            cache_dir=self._cache_dir if cache_dir is UNSET else cache_dir,  # type: ignore[arg-type]
            t_eval=self.t_eval, max_steps=self.max_steps)

    def _integrate(self, ode_system, y0, save_ts, params=None):
        # parameters in the state dtype (a float64 run uses float64 parameters)
        if params is not None:
            p = params.to(device=y0.device, dtype=y0.dtype)
        else:
            p = torch.tensor([float(v) for v in dict(ode_system.params).values()],
                             dtype=y0.dtype, device=y0.device)
        func = lambda t, y: ode_system.ode(t, y, p)  # noqa: E731
        with torch.no_grad():
            y = dopri5_shared(func, y0, float(self.t_span[0]), save_ts, self.rtol, self.atol,
                              self.max_steps)
        return save_ts, y
