"""CPU oracle for the Dormand-Prince basin-stability path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker.  The product package ``pybasin_amd`` never imports it.

Contents
--------
* ``solve``            ctypes front end of ``liboracle.so`` (dopri5_oracle.c):
                       torchdiffeq semantics per lane or with one shared dt,
                       or zigode semantics.
* ``save_grid``        ``Solver._prepare_tensors`` + TorchDiffEqSolver anchor
                       (src/pybasin/solvers/base.py:113-127,
                       src/pybasin/solvers/torchdiffeq_solver.py:122-131).
* ``expand_params``    IC-major expansion (src/pybasin/solvers/base.py:155-159).
* ``sample_uniform``   UniformRandomSampler (src/pybasin/sampler.py:61-65).
* ``filter_time_idx``  FeatureExtractor.filter_time (feature_extractor.py:85-130).
* ``features``         mean / std(ddof=0) / min / max / delta / log_delta
                       (ts_torch/calculators/torch_features_statistical.py:29-110)
                       with ``impute_extreme`` (torch_feature_utilities.py:60-94).
* ``sync_features``    SynchronizationFeatureExtractor._compute_sync_features
                       (case_studies/rossler_network/synchronization_feature_extractor.py:73-106).
* ``basin_stability``  _compute_bs / get_errors (basin_stability_estimator.py:586-625).
* ``knn1``             1-nearest-neighbour label (sklearn KNeighborsClassifier(1)).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SYS_DECAY, SYS_LINEAR2, SYS_PENDULUM, SYS_DUFFING, SYS_LORENZ, SYS_ROSSLER = range(6)
MODE_TD_LANE, MODE_TD_SHARED, MODE_ZIG = 0, 1, 2
_NPARAMS = {SYS_DECAY: 1, SYS_LINEAR2: 4, SYS_PENDULUM: 3, SYS_DUFFING: 3, SYS_LORENZ: 3, SYS_ROSSLER: 4}

_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_solve.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, dp, dp,
            ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_int64, dp,
            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
        ]
        L.or_solve.restype = ctypes.c_int
        L.or_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp,
                             ctypes.c_int, dp]
        L.or_rhs.restype = ctypes.c_int
        L.or_set_rossler_edges.argtypes = [ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        L.or_set_rossler_edges.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_rossler_edges(edges: np.ndarray) -> None:
    e = np.ascontiguousarray(edges, dtype=np.int32)
    ei = np.ascontiguousarray(e[:, 0])
    ej = np.ascontiguousarray(e[:, 1])
    rc = lib().or_set_rossler_edges(ei.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    ej.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(e))
    if rc != 0:
        raise ValueError("edge list must hold 400 edges")


def save_grid(t_span, t_steps, t_eval=None, dtype=np.float32):
    """Save times exactly as the reference builds them: torch.linspace in the
    state dtype over t_eval (or t_span), then promoted to float64 by
    torchdiffeq.  Returns (save_ts_in_dtype, save_ts_f64, t0)."""
    import torch

    start, end = t_eval if t_eval is not None else t_span
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    ts = torch.linspace(start, end, t_steps, dtype=tdt)
    return ts.numpy(), ts.numpy().astype(np.float64), float(t_span[0])


def solve(sys: int, y0: np.ndarray, params: np.ndarray, save_ts64: np.ndarray, t0: float,
          rtol: float = 1e-8, atol: float = 1e-6, bits: int = 32, mode: int = MODE_TD_LANE,
          t1: float | None = None, max_steps: int = 10**9, threads: int = 1):
    """Integrate rows of y0 (n, S).  params: (n_p,) shared or (n, n_p) per row.
    Returns (traj (n_save, n, S) float64, status (n,), nsteps (n,))."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    n, S = y0.shape
    params = np.asarray(params, dtype=np.float64)
    n_p = _NPARAMS[sys]
    if params.ndim == 1:
        params = np.ascontiguousarray(params.reshape(1, -1))
        ps = 0
    else:
        params = np.ascontiguousarray(params)
        ps = params.shape[1]
    assert params.shape[-1] == n_p
    save = np.ascontiguousarray(save_ts64, dtype=np.float64)
    out = np.zeros((len(save), n, S), dtype=np.float64)
    status = np.zeros(n, dtype=np.int32)
    nsteps = np.zeros(n, dtype=np.int64)
    if t1 is None:
        t1 = float(save[-1])
    rc = lib().or_solve(sys, bits, mode, n, S, _dp(y0), _dp(params), n_p, ps, t0, t1, _dp(save),
                        len(save), rtol, atol, max_steps, _dp(out),
                        status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                        nsteps.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), threads)
    if rc < 0:
        raise RuntimeError(f"oracle or_solve failed: {rc}")
    return out, status, nsteps


def rhs(sys: int, t: float, y, p, bits: int = 64) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    dy = np.zeros_like(y)
    rc = lib().or_rhs(sys, bits, len(y), t, _dp(y), _dp(p), len(p), _dp(dy))
    if rc != 0:
        raise ValueError("bad rhs arguments")
    return dy


def expand_params(y0: np.ndarray, params: np.ndarray):
    """IC-major: trajectory ic*P + p carries (y0[ic], params[p])."""
    B, P = y0.shape[0], params.shape[0]
    return np.repeat(y0, P, axis=0), np.tile(params, (B, 1))


def sample_uniform(n: int, lo, hi, seed: int) -> np.ndarray:
    """torch.manual_seed(seed); torch.rand(n, S) * (hi - lo) + lo in float32 on CPU."""
    import torch

    torch.manual_seed(seed)
    lo_t = torch.tensor(lo, dtype=torch.float32)
    hi_t = torch.tensor(hi, dtype=torch.float32)
    return (torch.rand(n, len(lo)) * (hi_t - lo_t) + lo_t).numpy()


def filter_time_idx(save_ts: np.ndarray, time_steady: float | None) -> int:
    """Index of the first save time >= time_steady (0 if none); default
    time_steady = t0 + 0.85 (t1 - t0) of the save grid."""
    if time_steady is None:
        t0, t1 = float(save_ts[0]), float(save_ts[-1])
        time_steady = t0 + 0.85 * (t1 - t0)
    idx = np.nonzero(save_ts >= time_steady)[0]
    return int(idx[0]) if len(idx) else 0


def impute_extreme(f: np.ndarray, extreme: float = 1e10) -> np.ndarray:
    f = f.copy()
    f[np.isposinf(f)] = extreme
    f[np.isneginf(f)] = -extreme
    f[np.isnan(f)] = extreme
    return f


def features(window: np.ndarray) -> dict[str, np.ndarray]:
    """Per-state streaming-feature set over a (N', B, S) window, in float64."""
    w = window.astype(np.float64)
    mean = w.mean(axis=0)
    std = w.std(axis=0)  # ddof = 0, as torch correction=0
    mx = w.max(axis=0)
    mn = w.min(axis=0)
    delta = np.abs(mx - mean)
    return {
        "mean": mean, "standard_deviation": std, "maximum": mx, "minimum": mn,
        "delta": delta, "log_delta": np.log(delta + 1e-10),
    }


def sync_features(final_state: np.ndarray, n_nodes: int) -> np.ndarray:
    x = final_state[:, :n_nodes]
    y = final_state[:, n_nodes:2 * n_nodes]
    z = final_state[:, 2 * n_nodes:]
    dx = x.max(1) - x.min(1)
    dy = y.max(1) - y.min(1)
    dz = z.max(1) - z.min(1)
    return np.stack([dx, dy, dz, np.maximum(np.maximum(dx, dy), dz)], axis=1)


def knn1(feat: np.ndarray, templates: np.ndarray, labels) -> np.ndarray:
    d = ((feat[:, None, :] - templates[None, :, :]) ** 2).sum(-1)
    return np.asarray(labels)[np.argmin(d, axis=1)]


def basin_stability(labels: np.ndarray) -> dict:
    """Fractions with binomial standard errors e_abs = sqrt(S(1-S)/n)."""
    labels = np.asarray(labels)
    n = len(labels)
    out = {}
    for lab in np.unique(labels):
        s = float(np.sum(labels == lab)) / n
        out[str(lab)] = (s, float(np.sqrt(s * (1 - s) / n)))
    return out
