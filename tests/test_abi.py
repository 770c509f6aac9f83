"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU compute)."""

import ctypes
import glob
import os
import re

import pytest

from pybasin_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bh_solve_batch" in syms and "bh_version" in syms
    assert set(syms) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} missing: run python -m pybasin_amd.build")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s


def test_version_and_argument_errors_without_gpu():
    L = _lib.lib()
    assert L.bh_version() >= 10000
    # invalid system id is rejected before any HIP call
    rc = L.bh_solve_batch(99, 0, None, 1, 2, None, 3, 0, 0.0, None, 1, 1e-8, 1e-6, 10, 0, 0.0, 0,
                          None, None, None, None, None, None, 0, 0, None)
    assert rc == -1
    assert b"system" in L.bh_last_error()
    # wrong state dimension for the pendulum
    rc = L.bh_solve_batch(2, 0, None, 1, 3, None, 3, 0, 0.0, None, 1, 1e-8, 1e-6, 10, 0, 0.0, 0,
                          None, None, None, None, None, None, 0, 0, None)
    assert rc == -5
    with pytest.raises(ValueError):
        _lib.check(rc, "bh_solve_batch")
    # the tuner checks its candidate lists before any HIP call
    rc = L.bh_autotune_lane(4, 0, None, 1, 3, None, 3, 0, 0.0, None, 1, 1e-8, 1e-6, 10, 0, 0.0, 0,
                            None, None, None, None, None, None, 0, None, 0, None, 0, None, 0, 1,
                            None, None, None, None, None)
    assert rc == -5 and b"candidate" in L.bh_last_error()
    # the two-phase entry checks its step budget and the workspace size query its system
    rc = L.bh_solve_batch_phased(4, 0, None, 1, 3, None, 3, 0, 0.0, None, 1, 1e-8, 1e-6, 10, 0, 0.0,
                                 0, None, None, None, None, None, 0, None, 0, None)
    assert rc == -5 and b"phase_iters" in L.bh_last_error()
    assert L.bh_solve_batch_phased_workspace_bytes(99, 0, 2, 10) < 0
    need = L.bh_solve_batch_phased_workspace_bytes(4, 0, 3, 1000)
    assert need > 256 + 8 * 1000
    # a workspace below that size is refused before any device work (the
    # pointers are never dereferenced on this path)
    fake = 4096
    rc = L.bh_solve_batch_phased(4, 0, fake, 1000, 3, fake, 3, 0, 0.0, fake, 10, 1e-8, 1e-6, 10, 0,
                                 0.0, 0, None, None, None, None, None, 5, fake, need - 1, None)
    assert rc == -5 and b"workspace" in L.bh_last_error()


def test_noalloc_entry_checks_size_before_device_work():
    """The allocation-free two-phase entry point: its size limit, and the
    errors it returns before touching a device."""
    L = _lib.lib()
    sid, dt = _lib.BH_SYS_DUFFING, _lib.BH_F32
    m = L.bh_solve_batch_phased_noalloc_max_n(sid, dt, 2)
    assert m > 300_000
    assert L.bh_solve_batch_phased_workspace_bytes(sid, dt, 2, m) <= _lib.BH_NOALLOC_SCRATCH_BYTES
    assert L.bh_solve_batch_phased_workspace_bytes(sid, dt, 2, m + 1) > _lib.BH_NOALLOC_SCRATCH_BYTES
    args = lambda n, k: (sid, dt, None, n, 2, None, 3, 0, 0.0, None, 10, 1e-8, 1e-6,  # noqa: E731
                         1000, 0, 0.0, 0, None, None, None, None, None, k, None)
    assert L.bh_solve_batch_phased_noalloc(*args(m + 1, 100)) == -5
    assert b"static" in L.bh_last_error()
    assert L.bh_solve_batch_phased_noalloc(*args(10, 0)) == -5  # phase_iters >= 1
    assert L.bh_solve_batch_phased_noalloc(*args(0, 100)) == 0  # nothing to do
    assert L.bh_solve_batch_phased_noalloc_max_n(99, dt, 2) < 0


def test_rossler_edges_entry_checks_the_graph_before_device_work():
    """bh_solve_rossler_edges validates its host edge arrays (node range, at
    most 12 entries per node) before anything reaches a device; n = 0 with a
    valid graph does nothing."""
    import numpy as np

    L = _lib.lib()

    def call(ei, ej, n=0):
        ei = np.ascontiguousarray(ei, dtype=np.int32)
        ej = np.ascontiguousarray(ej, dtype=np.int32)
        return L.bh_solve_rossler_edges(_lib.BH_F32, None, n, None, 0, ei.ctypes.data, ej.ctypes.data,
                                        len(ei), 0.0, 1.0, 1e-3, 1e-6, 1000, 0, 0.0, None, None, None,
                                        None, 0, 0, None)

    ring = np.arange(100)
    ei = np.concatenate([ring, (ring + 1) % 100])
    ej = np.concatenate([(ring + 1) % 100, ring])
    assert call(ei, ej) == 0
    assert call(np.append(ei, 100), np.append(ej, 0)) == -5
    assert b"outside" in L.bh_last_error()
    assert call(np.append(ei, 3), np.append(ej, -1)) == -5
    hub_i = np.zeros(13, dtype=np.int32)
    hub_j = np.arange(1, 14, dtype=np.int32)
    assert call(hub_i, hub_j) == -5
    assert b"more than 12" in L.bh_last_error()
    assert call(ei[:0], ej[:0]) == 0  # no edges: uncoupled nodes


def test_window_entries_check_arguments_before_device_work():
    """Every window-feature entry point refuses bad arguments with -5 and a
    message before any HIP call (the pointers below are never dereferenced)."""
    L = _lib.lib()
    fake, F32, F64 = 4096, _lib.BH_F32, _lib.BH_F64

    def refused(rc, word):
        assert rc == -5, rc
        assert word in L.bh_last_error(), L.bh_last_error()

    # common checks: empty window, dtype, null trajectory
    refused(L.bh_window_features(F32, fake, 10, 4, 2, 10, fake, None), b"empty window")
    refused(L.bh_window_spectrum(7, fake, 10, 4, 2, 0, 5, fake, None, None), b"dtype")
    refused(L.bh_window_autocov(F32, None, 10, 4, 2, 0, 3, fake, None), b"null trajectory")
    # per-entry limits
    refused(L.bh_window_spectrum(F32, fake, 10, 4, 2, 0, 5000, fake, None, None), b"n_coeff")
    refused(L.bh_window_spectrum(F32, fake, 10, 4, 2, 0, 0, None, None, None), b"nothing")
    refused(L.bh_window_spectrum(F32, fake, 3174, 4, 2, 0, 0, None, fake, None), b"LDS budget")
    refused(L.bh_window_spectrum(F64, fake, 7937, 4, 2, 0, 5, fake, None, None), b"LDS budget")
    refused(L.bh_window_peak_counts(F32, fake, 5291, 4, 2, 0, 10, 0, 0, 0, fake, None), b"at most 5290")
    refused(L.bh_window_peak_counts(F64, fake, 3969, 4, 2, 0, 10, 0, 0, 0, fake, None), b"at most 3968")
    refused(L.bh_window_embedding(F64, fake, 4097, 4, 2, 0, _lib.BH_EMB_LYAP_E, fake, None),
            b"LDS budget")
    refused(L.bh_window_autocov(F32, fake, 10, 4, 2, 0, 5000, fake, None), b"max_lag")
    refused(L.bh_window_sorted(F64, fake, 4097, 4, 2, 0, fake, None, None, None), b"LDS budget")
    refused(L.bh_window_sorted(F32, fake, 8193, 4, 2, 0, fake, None, None, None), b"LDS budget")
    refused(L.bh_window_sorted(F32, fake, 100, 4, 2, 0, None, None, None, None), b"out_dup")
    refused(L.bh_window_change_quantiles(F32, fake, 100, 4, 2, 0, None, fake, None), b"levels")
    refused(L.bh_window_trends(F32, fake, 100, 4, 2, 0, 1, None, None), b"null out")
    refused(L.bh_window_peak_counts(F32, fake, 100, 4, 2, 0, 10, 65, 0, 0, fake, None), b"supports")
    refused(L.bh_window_perm_entropy(F32, fake, 4053, 4, 2, 0, fake, None), b"LDS budget")
    refused(L.bh_window_embedding(F32, fake, 100, 4, 2, 0, 9, fake, None), b"which")
    bins = (ctypes.c_int32 * 2)(2, 300)
    refused(L.bh_window_lempel_ziv(F32, fake, 100, 4, 2, 0, ctypes.cast(bins, ctypes.c_void_p), 2,
                                   fake, None), b"bins")
    refused(L.bh_window_lempel_ziv(F32, fake, 100, 4, 2, 0, ctypes.cast(bins, ctypes.c_void_p), 9,
                                   fake, None), b"n_bins")
    ok_bins = (ctypes.c_int32 * 1)(2)
    refused(L.bh_window_lempel_ziv(F32, fake, 1025, 4, 2, 0, ctypes.cast(ok_bins, ctypes.c_void_p), 1,
                                   fake, None), b"1,024 rows")
    # n = 0: nothing to do, success without device work
    assert L.bh_window_trends(F32, fake, 100, 0, 2, 0, 1, fake, None) == 0
