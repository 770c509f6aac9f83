"""ctypes binding of libbasinhip.so (declared in include/basinhip.h).

The library is loaded from the package directory only.  There is no
fallback: if it is missing or fails to load, every HIP entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libbasinhip.so")

BH_SYS_DECAY, BH_SYS_LINEAR2, BH_SYS_PENDULUM, BH_SYS_DUFFING, BH_SYS_LORENZ, BH_SYS_ROSSLER100 = range(6)
BH_F32, BH_F64 = 0, 1
BH_LANE_OK, BH_LANE_MAX_STEPS, BH_LANE_DT_UNDERFLOW, BH_LANE_NONFINITE, BH_LANE_DIVERGED = range(5)
BH_FLAG_STOP = 0x1
BH_FLAG_INF_FILL = 0x2
BH_FLAG_COUNT_SAVE_STEPS = 0x4
BH_FLAG_TREND = 0x8
BH_REFILL_QUORUM_SHIFT = 8
BH_NFEAT = 4
BH_NFEAT_TREND = 5
BH_WORKSPACE_BYTES = 256
# slots of bh_window_features (include/basinhip.h BH_WF_*)
WF = dict(N=0, SUM=1, SUMSQ=2, MIN=3, MAX=4, FIRST_MAX=5, LAST_MAX=6, FIRST_MIN=7, LAST_MIN=8,
          ABS_CHANGE=9, SQ_CHANGE=10, X_FIRST=11, X_SECOND=12, X_SECOND_LAST=13, X_LAST=14,
          COUNT_POS=15, COUNT_NEG=16, COUNT_ZERO=17, COUNT_ONE=18, COUNT_MINUS_ONE=19,
          RANGE_UNIT=20, RANGE_NONPOS=21, RANGE_NONNEG=22, CROSS_0=23, CROSS_M1=24, CROSS_P1=25,
          CHUNK=26, C3=36, TREV=39, SUM_ABS=42, TOP_ABS=43, PEAKS_1=50, PEAKS_3=51, PEAKS_5=52,
          CM2=53, CM3=54, CM4=55, SS_TX=56, COUNT_ABOVE_MEAN=57, COUNT_BELOW_MEAN=58,
          STRIKE_ABOVE=59, STRIKE_BELOW=60, COUNT_MAX=61, COUNT_MIN=62, BEYOND=63, AUTOCOV=73,
          MASS=82, BINS=90)
BH_WF_SLOTS = 100
BH_EMB_CORR_DIM, BH_EMB_LYAP_R, BH_EMB_LYAP_E = 0, 1, 2
BH_NOALLOC_SCRATCH_BYTES = 64 << 20

# every symbol include/basinhip.h declares
EXPORTS = ("bh_version", "bh_last_error", "bh_solve_batch", "bh_solve_batch_peaks",
           "bh_solve_batch_phased_workspace_bytes", "bh_solve_batch_phased",
           "bh_solve_batch_phased_noalloc", "bh_solve_batch_phased_noalloc_max_n", "bh_autotune_lane", "bh_solve_rossler", "bh_solve_rossler_edges", "bh_gemm_bf16",
           "bh_mlp_bias_act", "bh_mlp_silu_bwd", "bh_mlp_softmax_ce", "bh_mlp_colsum",
           "bh_mlp_adam", "bh_mlp_prep_input", "bh_mlp_classify", "bh_mlp_infer",
           "bh_mlp_train_workspace_bytes", "bh_mlp_train_step", "bh_knn",
           "bh_hdbscan_workspace_bytes", "bh_core_distances", "bh_hdbscan_mst", "bh_hdbscan_labels",
           "bh_silhouette",
           "bh_window_order_stats", "bh_window_autocorr_periodicity", "bh_window_spectral_ratio",
           "bh_window_features", "bh_window_spectrum", "bh_window_sorted",
           "bh_window_change_quantiles", "bh_window_autocov", "bh_window_trends",
           "bh_window_peak_counts", "bh_window_perm_entropy", "bh_window_embedding",
           "bh_window_lempel_ziv")

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


class HipLibraryError(RuntimeError):
    pass


def _declare(L: ctypes.CDLL) -> None:
    vp, i32, i64, f64, u32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32
    L.bh_version.restype = i32
    L.bh_version.argtypes = []
    L.bh_last_error.restype = ctypes.c_char_p
    L.bh_last_error.argtypes = []
    L.bh_solve_batch.restype = i32
    L.bh_solve_batch.argtypes = [i32, i32, vp, i64, i32, vp, i32, i64, f64, vp, i32, f64, f64,
                                 i64, u32, f64, i32, vp, vp, vp, vp, vp, vp, i64, i32, vp]
    L.bh_solve_batch_peaks.restype = i32
    L.bh_solve_batch_peaks.argtypes = [i32, i32, vp, i64, i32, vp, i32, i64, f64, vp, i32, f64,
                                       f64, i64, u32, f64, i32, vp, vp, vp, vp, vp, i32, vp, vp,
                                       vp, i64, i32, vp]
    L.bh_solve_batch_phased_workspace_bytes.restype = i64
    L.bh_solve_batch_phased_workspace_bytes.argtypes = [i32, i32, i32, i64]
    L.bh_solve_batch_phased.restype = i32
    L.bh_solve_batch_phased.argtypes = [i32, i32, vp, i64, i32, vp, i32, i64, f64, vp, i32, f64,
                                        f64, i64, u32, f64, i32, vp, vp, vp, vp, vp, i32, vp, i64,
                                        vp]
    L.bh_solve_batch_phased_noalloc.restype = i32
    L.bh_solve_batch_phased_noalloc.argtypes = [i32, i32, vp, i64, i32, vp, i32, i64, f64, vp, i32,
                                                f64, f64, i64, u32, f64, i32, vp, vp, vp, vp, vp,
                                                i32, vp]
    L.bh_solve_batch_phased_noalloc_max_n.restype = i64
    L.bh_solve_batch_phased_noalloc_max_n.argtypes = [i32, i32, i32]
    L.bh_autotune_lane.restype = i32
    L.bh_autotune_lane.argtypes = [i32, i32, vp, i64, i32, vp, i32, i64, f64, vp, i32, f64, f64,
                                   i64, u32, f64, i32, vp, vp, vp, vp, vp, vp, i64, vp, i32, vp,
                                   i32, vp, i32, i32, vp, vp, vp, vp, vp]
    L.bh_solve_rossler.restype = i32
    L.bh_solve_rossler.argtypes = [i32, vp, i64, vp, i64, vp, f64, f64, f64, f64, i64, u32,
                                   f64, vp, vp, vp, vp, i64, i32, vp]
    L.bh_solve_rossler_edges.restype = i32
    L.bh_solve_rossler_edges.argtypes = [i32, vp, i64, vp, i64, vp, vp, i64, f64, f64, f64, f64, i64,
                                         u32, f64, vp, vp, vp, vp, i64, i32, vp]
    L.bh_hdbscan_workspace_bytes.restype = i64
    L.bh_hdbscan_workspace_bytes.argtypes = [i64]
    L.bh_mlp_train_workspace_bytes.restype = i64
    L.bh_mlp_train_workspace_bytes.argtypes = [i64, i32]
    f32 = ctypes.c_float
    sigs = {
        "bh_gemm_bf16": [vp, i64, i64, vp, i64, i64, vp, i64, i64, i64, i64, i32, vp],
        "bh_mlp_bias_act": [vp, i64, i64, vp, i32, vp, vp, i64, vp, vp],
        "bh_mlp_silu_bwd": [vp, vp, i64, vp, vp, vp],
        "bh_mlp_softmax_ce": [vp, vp, i64, vp, vp, vp, vp],
        "bh_mlp_colsum": [vp, i64, i64, vp, vp],
        "bh_mlp_adam": [vp, vp, vp, vp, i64, f32, f32, f32, f32, i64, vp, vp],
        "bh_mlp_prep_input": [vp, i64, f32, f32, f32, f32, f32, f32, vp, vp],
        "bh_mlp_classify": [vp, vp, i64, vp, vp, vp],
        "bh_mlp_infer": [vp, i64, vp, vp, f32, f32, f32, f32, f32, f32, vp, vp, vp],
        "bh_mlp_train_step": [vp, vp, i64, vp, i32, vp, vp, vp, vp, vp, vp, vp, f32, f32, f32, f32,
                              i64, vp, vp, i64, vp],
        "bh_knn": [vp, i64, vp, i64, i32, i32, vp, vp, vp],
        "bh_core_distances": [vp, i64, i32, i32, vp, vp, i64, vp],
        "bh_hdbscan_mst": [vp, i64, i32, vp, vp, vp, vp, vp, i64, vp],
        "bh_hdbscan_labels": [vp, vp, vp, i64, vp, i32, vp, vp],
        "bh_silhouette": [vp, i64, i32, vp, vp, i32, vp, vp, vp, vp, vp],
        "bh_window_order_stats": [i32, vp, i64, i64, i32, i64, vp, i32, vp, vp],
        "bh_window_autocorr_periodicity": [i32, vp, i64, i64, i32, i64, i32, f64, vp, vp, vp],
        "bh_window_spectral_ratio": [i32, vp, i64, i64, i32, i64, vp, vp],
        "bh_window_features": [i32, vp, i64, i64, i32, i64, vp, vp],
        "bh_window_spectrum": [i32, vp, i64, i64, i32, i64, i32, vp, vp, vp],
        "bh_window_sorted": [i32, vp, i64, i64, i32, i64, vp, vp, vp, vp],
        "bh_window_change_quantiles": [i32, vp, i64, i64, i32, i64, vp, vp, vp],
        "bh_window_autocov": [i32, vp, i64, i64, i32, i64, i32, vp, vp],
        "bh_window_trends": [i32, vp, i64, i64, i32, i64, i32, vp, vp],
        "bh_window_peak_counts": [i32, vp, i64, i64, i32, i64, i32, i32, i32, i32, vp, vp],
        "bh_window_perm_entropy": [i32, vp, i64, i64, i32, i64, vp, vp],
        "bh_window_embedding": [i32, vp, i64, i64, i32, i64, i32, vp, vp],
        "bh_window_lempel_ziv": [i32, vp, i64, i64, i32, i64, vp, i32, vp, vp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.restype = i32
        fn.argtypes = args


def lib() -> ctypes.CDLL:
    """Load libbasinhip.so once; raise HipLibraryError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise HipLibraryError(
                    f"{LIB_PATH} is not built; run `python -m pybasin_amd.build` "
                    "(the HIP path has no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            _declare(L)
            _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().bh_last_error().decode(errors="replace")
        if rc == -5:
            raise ValueError(f"{what}: {msg}")
        raise HipLibraryError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()
