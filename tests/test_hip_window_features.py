"""GPU parity of the dynamical-system feature set (ts_torch/settings.py:407-417).

* amplitude and the linear-trend slope are streaming statistics of the
  integration kernel (the slope under BH_FLAG_TREND): equal to the same
  statistics of the kernel's own kept window;
* autocorrelation_periodicity and spectral_frequency_ratio are series
  kernels over a kept window (bh_window_autocorr_periodicity,
  bh_window_spectral_ratio): against the oracle's restatements (the
  reference's FFT formulation, pinned by closed forms) on integrated
  Duffing windows and on closed-form signals.
"""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pybasin_amd import _lib
from pybasin_amd.features import (DYNAMICAL_SYSTEM_FEATURES, HipFusedFeatureExtractor, Solution,
                                  StatFeatureExtractor, stats_from_window, window_feature)
from pybasin_amd.solvers import HipDopri5Solver
from pybasin_amd.systems import DuffingODE, LorenzODE

pytestmark = pytest.mark.gpu
DUFF = {"delta": 0.08, "k3": 1.0, "A": 0.2}


def _duffing(dtype, n=2000, seed=9):
    y0 = torch.tensor(O.sample_uniform(n, [-1.0, -0.5], [1.0, 1.0], seed=seed), dtype=dtype).cuda()
    solver = HipDopri5Solver(t_span=(0, 400), t_steps=1000, t_eval=(200.0, 400.0), device="cuda",
                             cache_dir=None)
    return solver.integrate_fused(DuffingODE(DUFF), y0, time_steady=300.0, keep_trajectories=True,
                                  trend=True)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_slope_and_amplitude_equal_kept_window(cuda_device, dtype):
    res = _duffing(dtype)
    assert res.stats.shape[-1] == _lib.BH_NFEAT_TREND
    w = res.y[res.steady_idx:]
    ref = stats_from_window(w, trend=True)
    scale = w.double().abs().amax(0) / w.shape[0]
    np.testing.assert_allclose(res.stats[..., 4].cpu().numpy(), ref[..., 4].cpu().numpy(), rtol=1e-7,
                               atol=float(1e-9 * scale.max()))
    torch.testing.assert_close(res.stats[..., :4], ref[..., :4], rtol=1e-9, atol=1e-9)


def test_fused_slope_of_stopped_lanes(cuda_device):
    """Lanes stopped by the divergence test fill the rest of the window at
    once; the slope's position sums cover those samples too."""
    y0 = torch.tensor(O.sample_uniform(512, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], seed=4)).cuda()
    solver = HipDopri5Solver(t_span=(0, 60), t_steps=300, t_eval=(20.0, 60.0), device="cuda",
                             cache_dir=None, stop_threshold=200.0)
    res = solver.integrate_fused(LorenzODE({"sigma": 0.12, "r": 0.0, "b": -0.6}), y0,
                                 time_steady=30.0, keep_trajectories=True, trend=True)
    assert (res.status == _lib.BH_LANE_DIVERGED).any() and (res.status == 0).any()
    ref = stats_from_window(res.y[res.steady_idx:], trend=True)
    torch.testing.assert_close(res.stats[..., 4], ref[..., 4], rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_series_kernels_match_oracle_on_duffing_windows(cuda_device, dtype):
    res = _duffing(dtype)
    si = res.steady_idx
    w = res.y[si:].double().cpu().numpy()
    np_dt = np.float32 if dtype == torch.float32 else np.float64
    strength, period = O.autocorrelation_periodicity(w)
    # outputs are in the window's dtype
    ratio = O.spectral_frequency_ratio(w).astype(np_dt).astype(np.float64)
    g_s = window_feature("autocorrelation_periodicity", {"output": "strength"}, res.y, si).double().cpu().numpy()
    g_p = window_feature("autocorrelation_periodicity", {"output": "period"}, res.y, si).double().cpu().numpy()
    g_r = window_feature("spectral_frequency_ratio", None, res.y, si).double().cpu().numpy()
    same_p = g_p == period
    same_r = g_r == ratio
    print(f"[result] {dtype}: periods equal {same_p.mean():.5f}, ratios equal {same_r.mean():.5f}, "
          f"max strength diff where periods agree {np.abs(g_s - strength)[same_p].max():.2e}; "
          f"periods seen {sorted(set(period.ravel().tolist()))[:8]}", flush=True)
    assert same_p.mean() >= 0.999 and same_r.mean() >= 0.995
    tol = 1e-4 if dtype == torch.float32 else 1e-9
    assert np.abs(g_s - strength)[same_p].max() <= tol


def test_series_kernels_closed_forms(cuda_device):
    n = 1000
    i = torch.arange(n, dtype=torch.float64)
    sig = torch.stack([torch.sin(2 * math.pi * i / 50),
                       torch.sin(2 * math.pi * 10 * i / n) + 0.5 * torch.sin(2 * math.pi * 30 * i / n),
                       torch.sin(2 * math.pi * 20 * i / n) + 0.6 * torch.sin(2 * math.pi * 10 * i / n),
                       torch.full((n,), 4.0)], 1)[:, :, None]  # (n, 4, 1)
    for dtype in (torch.float32, torch.float64):
        y = sig.to(dtype).cuda()
        per = window_feature("autocorrelation_periodicity", {"output": "period"}, y, 0)[:, 0].cpu()
        st = window_feature("autocorrelation_periodicity", {"output": "strength"}, y, 0)[:, 0].cpu()
        rat = window_feature("spectral_frequency_ratio", None, y, 0)[:, 0].cpu()
        assert per[0] == 50 and abs(float(st[0]) - 0.95) < 5e-3
        assert per[3] == 0 and st[3] == 0  # constant: no peak
        assert rat[1] == 3.0 and rat[2] == 0.5 and rat[3] == 0.0


def test_dynamical_system_set_fused_equals_cpu(cuda_device):
    res = _duffing(torch.float32, n=600, seed=2)
    fx = HipFusedFeatureExtractor(time_steady=300.0, features=DYNAMICAL_SYSTEM_FEATURES)
    assert fx.needs_window and fx.needs_trend
    got = fx.extract_from_result(res).cpu()
    cpu = StatFeatureExtractor(time_steady=300.0, features=DYNAMICAL_SYSTEM_FEATURES)
    y = res.y.cpu()
    want = cpu.extract_features(Solution(y[0], res.t.cpu(), y))
    assert fx.feature_names == cpu.feature_names and len(fx.feature_names) == 14
    for j, name in enumerate(fx.feature_names):
        if name.endswith("output_period") or name.endswith("spectral_frequency_ratio"):
            assert (got[:, j] == want[:, j]).float().mean() >= 0.995, name
        else:
            torch.testing.assert_close(got[:, j], want[:, j], rtol=2e-4, atol=2e-5, msg=name)


def test_series_window_limit_is_an_error(cuda_device):
    y = torch.zeros((20_000, 2, 1), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="LDS budget"):
        window_feature("spectral_frequency_ratio", None, y, 0)


# ---- comprehensive window features (bh_window_features) ---------------------

def _synthetic_windows():
    rng = np.random.default_rng(21)
    yield "normal", rng.normal(size=(300, 64, 2))
    yield "ties", np.round(rng.normal(size=(160, 40, 3)) * 2) / 2
    yield "constant", np.full((80, 5, 2), -1.0)
    yield "ramp", np.arange(50, dtype=float)[:, None, None] * np.ones((1, 3, 2)) * 0.1 - 2
    for n in (1, 2, 4, 7, 11, 12, 30):
        yield f"short{n}", rng.normal(size=(n, 9, 2))
    bad = rng.normal(size=(60, 6, 2))
    bad[5, 1, 0] = np.nan
    bad[7, 3, 1] = np.inf
    bad[0, 4, 0] = -np.inf
    yield "nonfinite", bad


def _check_against_oracle(blk, w64, start, tag, series_dtype=np.float64):
    from pybasin_amd.window_features import COMPREHENSIVE_FEATURES, comprehensive_feature

    x = w64[start:]
    N, B, S = x.shape
    med = torch.tensor(np.sort(x, axis=0)[(N - 1) // 2])
    bad = []
    from pybasin_amd.window_features import feature_family

    for name, grid in COMPREHENSIVE_FEATURES.items():
        for p in (grid or [None]):
            if feature_family(name, p) != "block":
                continue  # the other window kernels, checked below
            if name in ("skewness", "kurtosis", "augmented_dickey_fuller") and N <= 3:
                continue  # noise amplified by the reference's 1e-10 guards (and n - 3 = 0)
            got = comprehensive_feature(name, p, blk, N, med).numpy()
            for b in range(B):
                for s in range(S):
                    if name == "augmented_dickey_fuller" and np.all(np.isfinite(x[:, b, s])):
                        dd = np.diff(x[:, b, s])
                        if dd.var() <= 1e-12 * (dd * dd).mean():
                            # constant changes (a ramp): the t statistic is a ratio
                            # of two rounding residues on either side
                            continue
                    want = O.comprehensive(x[:, b, s], name, p, series_dtype)
                    g = float(got[b, s])
                    # pvalue and stderr of a perfect linear fit (1 - r^2 and the
                    # residuals at rounding level) are rounding noise of size ~1e-8
                    tol = 1e-6 if name.startswith("linear_trend") and p["attr"] in ("pvalue", "stderr") \
                        else 1e-11
                    if name == "partial_autocorrelation":
                        # the recursion divides by 1 - sum phi_j r_j, small for nearly
                        # periodic windows: summation-order differences of the
                        # autocorrelations (1e-16) grow to ~1e-8 there
                        tol = 1e-6
                    if not math.isfinite(want):
                        ok = not math.isfinite(g)
                    else:
                        ok = g == pytest.approx(want, rel=1e-9, abs=tol)
                    if not ok:
                        bad.append((tag, name, p, b, s, g, want))
                        assert len(bad) < 3, bad
    assert not bad, bad


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_features_kernel_equals_oracle_synthetic(cuda_device, dtype):
    from pybasin_amd.window_features import window_feature_block

    for tag, w in _synthetic_windows():
        wt = torch.tensor(w, dtype=dtype)
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            blk = window_feature_block(wt.cuda(), start).cpu()
            _check_against_oracle(blk, wt.double().numpy(), start, f"{tag}/{start}",
                                  np.float32 if dtype == torch.float32 else np.float64)
            # the slots of the CPU path (same definitions, another summation
            # order): integer-valued slots equal
            from pybasin_amd.window_features import window_feature_block as wfb
            cpu = wfb(wt, start)
            ints = [k for k in range(_lib.BH_WF_SLOTS)
                    if k not in (1, 2, 9, 10, 42, 53, 54, 55, 56) and not 26 <= k <= 41
                    and not 73 <= k <= 81]
            assert torch.equal(torch.nan_to_num(blk[ints], 7.0), torch.nan_to_num(cpu[ints], 7.0)), tag


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_features_on_integrated_duffing(cuda_device, dtype):
    """All comprehensive features of 2,000 integrated Duffing windows (the
    kept trajectory of a fused launch, window t in [300, 400]) against the
    oracle; the extractor's columns equal the derivation."""
    from pybasin_amd.window_features import window_feature_block

    res = _duffing(dtype)
    blk = window_feature_block(res.y, res.steady_idx).cpu()
    _check_against_oracle(blk, res.y.double().cpu().numpy(), res.steady_idx, "duffing",
                          np.float32 if dtype == torch.float32 else np.float64)
    fx = StatFeatureExtractor(time_steady=300.0, features={"kurtosis": None, "number_peaks": [{"n": 3}],
                                                          "linear_trend": [{"attr": "rvalue"}]})
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    want = np.array([[O.comprehensive(x[:, b, s], n_, p_) for s in range(2)
                      for n_, p_ in (("kurtosis", None), ("number_peaks", {"n": 3}),
                                     ("linear_trend", {"attr": "rvalue"}))] for b in range(x.shape[1])])
    np.testing.assert_allclose(f, want.astype(np.float32), rtol=2e-6, atol=1e-6)


def _check_spectrum(coeff, agg, x, tag, n_coeff, with_agg):
    """bh_window_spectrum's outputs against the oracle's DFT sums, series by
    series: bins to 1e-9 of the series' sum |x|, angles (where the bin is
    above 1e-6 of it) to 1e-6 rad, the aggregates to 1e-9."""
    N, B, S = x.shape
    nb = N // 2 + 1
    ks = sorted({k for k in (0, 1, 2, 3, 5, 8, 13, 31, 64, 99, nb - 1, nb, nb + 1) if k < n_coeff})
    for b in range(B):
        for s in range(S):
            xs = x[:, b, s]
            finite = np.all(np.isfinite(xs))
            scale = max(np.abs(xs).sum(), 1e-300) if finite else 1.0
            for k in ks:
                re, im = float(coeff[k, 0, b, s]), float(coeff[k, 1, b, s])
                if not finite:
                    if k < nb:
                        assert not (math.isfinite(re) and math.isfinite(im)), (tag, k, b, s)
                    continue
                want_re = O.comprehensive(xs, "fft_coefficient", {"coeff": k, "attr": "real"})
                want_im = O.comprehensive(xs, "fft_coefficient", {"coeff": k, "attr": "imag"})
                assert abs(re - want_re) <= 1e-9 * scale and abs(im - want_im) <= 1e-9 * scale, \
                    (tag, k, b, s, re, im, want_re, want_im)
                if k >= nb:
                    assert re == 0.0 and im == 0.0, (tag, k)
                if math.hypot(want_re, want_im) > 1e-6 * scale:
                    d = abs(math.atan2(im, re) - math.atan2(want_im, want_re))
                    assert min(d, 2 * math.pi - d) <= 1e-6, (tag, k, b, s)
                if k in (0, nb - 1) and (k == 0 or N % 2 == 0):
                    assert im == 0.0 and math.copysign(1.0, im) > 0, (tag, k)  # real bins, +0
            if with_agg:
                for j, (name, p) in enumerate((("fft_aggregated", {"aggtype": "centroid"}),
                                               ("fft_aggregated", {"aggtype": "variance"}),
                                               ("fft_aggregated", {"aggtype": "skew"}),
                                               ("fft_aggregated", {"aggtype": "kurtosis"}),
                                               ("fourier_entropy", {"bins": 10}))):
                    want = O.comprehensive(xs, name, p)
                    g = float(agg[j, b, s])
                    if not math.isfinite(want):
                        assert not math.isfinite(g), (tag, name, p, b, s)
                    else:
                        assert g == pytest.approx(want, rel=1e-9, abs=1e-9), (tag, name, p, b, s, g,
                                                                              want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_spectrum_kernel_equals_oracle_synthetic(cuda_device, dtype):
    """Bins 0..99 alone (one or two bins per lane), bins past the last one
    (zeros), the full spectrum with its aggregates (up to four bins per lane
    in several rounds), the aggregates alone; ragged, short, constant and
    non-finite windows; windows starting at row 0 and 1."""
    from pybasin_amd.window_features import window_spectrum

    for tag, w in _synthetic_windows():
        wt = torch.tensor(w, dtype=dtype)
        x = wt.double().numpy()
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            nb = (w.shape[0] - start) // 2 + 1
            for n_coeff, with_agg in ((100, False), (nb + 2, True), (0, True), (40, False)):
                coeff, agg = window_spectrum(wt.cuda(), start, n_coeff, with_agg)
                if n_coeff:
                    assert coeff.shape == (n_coeff, 2) + w.shape[1:]
                else:
                    assert coeff is None
                    coeff = torch.zeros((0, 2) + w.shape[1:], dtype=torch.float64)
                _check_spectrum(coeff.cpu(), agg.cpu() if with_agg else None, x[start:],
                                f"{tag}/{start}/{n_coeff}/{with_agg}", n_coeff, with_agg)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_spectrum_on_integrated_duffing(cuda_device, dtype):
    """The rfft of 2,000 integrated Duffing windows (the kept trajectory of a
    fused launch) equals numpy's FFT on every series (bins 0..99 and the
    aggregates formed from numpy's full spectrum), the oracle on eight; the
    extractor's spectrum columns equal the kernel's outputs."""
    from pybasin_amd.window_features import spectrum_feature, window_spectrum

    res = _duffing(dtype)
    coeff, agg = window_spectrum(res.y, res.steady_idx, 100, True)
    coeff, agg = coeff.cpu().numpy(), agg.cpu().numpy()
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    X = np.fft.rfft(x, axis=0)
    scale = np.abs(x).sum(0)
    assert np.all(np.abs(coeff[:, 0] - X[:100].real) <= 1e-9 * scale)
    assert np.all(np.abs(coeff[:, 1] - X[:100].imag) <= 1e-9 * scale)
    P = X.real ** 2 + X.imag ** 2
    pn = P / (P.sum(0) + 1e-10)
    bins = np.arange(P.shape[0], dtype=float)[:, None, None]
    cen = (bins * pn).sum(0)
    np.testing.assert_allclose(agg[0], cen, rtol=1e-9)
    np.testing.assert_allclose(agg[1], ((bins - cen) ** 2 * pn).sum(0), rtol=1e-9)
    pc = np.maximum(pn, 1e-10)
    np.testing.assert_allclose(agg[4], -(pc * np.log(pc)).sum(0), rtol=1e-9)
    _check_spectrum(coeff[:, :, :4], agg[:, :4], x[:, :4], "duffing", 100, True)
    feats = {"fft_coefficient": [{"coeff": 7, "attr": "angle"}], "fft_aggregated": [{"aggtype": "kurtosis"}],
             "spkt_welch_density": [{"coeff": 5}]}
    fx = StatFeatureExtractor(time_steady=300.0, features=feats)
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    c, a = torch.tensor(coeff), torch.tensor(agg)
    n = x.shape[0]
    want = np.stack([spectrum_feature(nm, p, c, a, n).numpy()[:, s] for s in range(2)
                     for nm, ps in feats.items() for p in ps], 1)
    np.testing.assert_array_equal(f, want.astype(np.float32))


def _check_sorted(dup, cq, x, tag, series_dtype):
    """bh_window_sorted's features against the oracle, series by series: the
    counts exactly, sums and change statistics to rel 1e-9."""
    from pybasin_amd.window_features import SORTED_FEATURES, sorted_feature

    N, B, S = x.shape
    for name, grid in SORTED_FEATURES.items():
        for p in (grid or [None]):
            got = sorted_feature(name, p, dup, cq, N).numpy()
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(x[:, b, s], name, p, series_dtype)
                    g = float(got[b, s])
                    if not math.isfinite(want):
                        assert not math.isfinite(g), (tag, name, p, b, s)
                    else:
                        assert g == pytest.approx(want, rel=1e-9, abs=1e-12), (tag, name, p, b, s,
                                                                               g, want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_sorted_kernel_equals_oracle_synthetic(cuda_device, dtype):
    """Duplicate / reoccurrence counts and the 60 change quantiles of the
    bitonic-sorted windows: ties, constants, ramps, short and non-finite
    windows, windows of 2,048 (float32) and 1,024 (float64) rows; the counts
    of the CPU path equal the kernel's."""
    from pybasin_amd.window_features import window_sorted

    rng = np.random.default_rng(31)
    big = 2048 if dtype == torch.float32 else 1024
    cases = list(_synthetic_windows()) + [("big", np.round(rng.normal(size=(big, 7, 2)) * 8) / 8)]
    for tag, w in cases:
        wt = torch.tensor(w, dtype=dtype)
        sd = np.float32 if dtype == torch.float32 else np.float64
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            dup, cq, med = window_sorted(wt.cuda(), start, True)
            dup, cq, med = dup.cpu(), cq.cpu(), med.cpu()
            _check_sorted(dup, cq, wt[start:].numpy().astype(sd), f"{tag}/{start}", sd)
            cdup, _, cmed = window_sorted(wt, start, False)
            assert torch.equal(torch.nan_to_num(dup[:3], 7.0), torch.nan_to_num(cdup[:3], 7.0)), tag
            # the lower middle value, as torch.median (NaN for a non-finite window)
            assert torch.equal(torch.nan_to_num(med, 7.0), torch.nan_to_num(cmed, 7.0)), tag
            fin = torch.isfinite(wt[start:]).all(0)
            assert torch.equal(med[fin], wt[start:].median(0).values[fin]), tag
            dup2, none, _ = window_sorted(wt.cuda(), start, False)
            assert none is None and torch.equal(torch.nan_to_num(dup2.cpu(), 7.0),
                                                torch.nan_to_num(dup, 7.0)), tag


def test_window_sorted_limit_is_an_error(cuda_device):
    from pybasin_amd.window_features import window_sorted

    with pytest.raises(ValueError, match="LDS budget"):
        window_sorted(torch.zeros(4097, 4, 1, dtype=torch.float64, device="cuda"), 0, True)
    with pytest.raises(ValueError, match="LDS budget"):
        window_sorted(torch.zeros(8193, 4, 1, dtype=torch.float32, device="cuda"), 0, False)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_sorted_long_windows_block_sort(cuda_device, dtype):
    """Windows past one wave's registers (float32 2,049..8,192, float64
    1,025..4,096 rows) are sorted by a whole block: the same statistics as
    the oracle, ties and a non-finite window included."""
    from pybasin_amd.window_features import window_sorted

    rng = np.random.default_rng(61)
    sd = np.float32 if dtype == torch.float32 else np.float64
    for N in ((2049, 5000, 8192) if dtype == torch.float32 else (1025, 3000, 4096)):
        w = np.round(rng.normal(size=(N, 3, 2)) * 16) / 16
        w[7, 2, 1] = np.nan
        wt = torch.tensor(w, dtype=dtype)
        dup, cq, med = window_sorted(wt.cuda(), 0, True)
        dup, cq, med = dup.cpu(), cq.cpu(), med.cpu()
        _check_sorted(dup, cq, wt.numpy().astype(sd), f"long{N}", sd)
        cdup, ccq, cmed = window_sorted(wt, 0, True)
        assert torch.equal(torch.nan_to_num(dup[:3], 7.0), torch.nan_to_num(cdup[:3], 7.0)), N
        assert torch.equal(torch.nan_to_num(med, 7.0), torch.nan_to_num(cmed, 7.0)), N


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_sorted_on_integrated_duffing(cuda_device, dtype):
    """Sorted-window features of 2,000 integrated Duffing windows against the
    oracle (on 64 series) and the CPU path's counts (on all); the extractor's
    columns equal the kernel's."""
    from pybasin_amd.window_features import sorted_feature, window_sorted

    res = _duffing(dtype)
    dup, cq, _ = window_sorted(res.y, res.steady_idx, True)
    dup, cq = dup.cpu(), cq.cpu()
    sd = np.float32 if dtype == torch.float32 else np.float64
    x = res.y.cpu().numpy()[res.steady_idx:]
    _check_sorted(dup[:, :32], cq[:, :32], x[:, :32], "duffing", sd)
    cdup, ccq, _ = window_sorted(res.y.cpu(), res.steady_idx, True)
    assert torch.equal(dup[:3], cdup[:3])
    np.testing.assert_allclose(cq.numpy(), ccq.numpy(), rtol=1e-9, atol=1e-12)
    feats = {"change_quantiles": [{"ql": 0.2, "qh": 0.8, "isabs": True, "f_agg": "var"}],
             "ratio_value_number_to_time_series_length": None}
    fx = StatFeatureExtractor(time_steady=300.0, features=feats)
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    n = x.shape[0]
    want = np.stack([sorted_feature(nm, p, dup, cq, n).numpy()[:, s] for s in range(2)
                     for nm, ps in feats.items() for p in (ps or [None])], 1)
    np.testing.assert_array_equal(f, want.astype(np.float32))


def _check_autocov(C, x, tag, max_lag):
    """bh_window_autocov against direct centred sums, and its features
    against the oracle (ar_coefficient: the Durbin-Levinson recursion against
    the oracle's LU solve, 1e-7)."""
    from pybasin_amd.window_features import AUTOCOV_FEATURES, autocov_feature

    N, B, S = x.shape
    memo: dict = {}
    xc = x - x.mean(0)
    scale = (xc * xc).sum(0)
    for k in range(max_lag + 1):
        want = (xc[: N - k] * xc[k:]).sum(0) if k < N else np.zeros((B, S))
        ok = np.isfinite(want) & np.isfinite(scale)
        assert np.all(np.abs(C[k][ok] - want[ok]) <= 1e-12 * (scale[ok] + 1)), (tag, k)
        if k < N:
            assert np.all(~np.isfinite(C[k][~ok])), (tag, k)
    for name, grid in AUTOCOV_FEATURES.items():
        for p in grid:
            got = autocov_feature(name, p, torch.tensor(C), N, memo).numpy()
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(x[:, b, s], name, p)
                    g = float(got[b, s])
                    if not math.isfinite(want):
                        assert not math.isfinite(g), (tag, name, p, b, s)
                        continue
                    tol = 1e-7 if name == "ar_coefficient" else 1e-11
                    assert g == pytest.approx(want, rel=1e-7, abs=tol), (tag, name, p, b, s, g, want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_autocov_kernel_equals_oracle_synthetic(cuda_device, dtype):
    """Autocovariances of lags 0..40 and 0..70 (two rounds of 64 lanes), lags
    past short windows (zeros), non-finite windows; agg_autocorrelation and
    ar_coefficient from them equal the oracle."""
    from pybasin_amd.window_features import window_autocov

    for tag, w in _synthetic_windows():
        wt = torch.tensor(w, dtype=dtype)
        x = wt.double().numpy()
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            for max_lag in (40, 70):
                C = window_autocov(wt.cuda(), start, max_lag).cpu().numpy()
                _check_autocov(C, x[start:], f"{tag}/{start}/{max_lag}", max_lag)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_autocov_on_integrated_duffing(cuda_device, dtype):
    """The autocovariance features of integrated Duffing windows against the
    oracle (24 series) and direct sums (all); the extractor's columns equal the
    kernel's, augmented_dickey_fuller included."""
    from pybasin_amd.window_features import autocov_feature, window_autocov

    res = _duffing(dtype)
    C = window_autocov(res.y, res.steady_idx, 40).cpu().numpy()
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    _check_autocov(C[:, :12], x[:, :12], "duffing", 40)
    feats = {"agg_autocorrelation": [{"f_agg": "median", "maxlag": 40}],
             "ar_coefficient": [{"coeff": 3, "k": 10}],
             "augmented_dickey_fuller": [{"attr": "teststat"}]}
    fx = StatFeatureExtractor(time_steady=300.0, features=feats)
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    n = x.shape[0]
    Ct = torch.tensor(C)
    want_agg = autocov_feature("agg_autocorrelation", feats["agg_autocorrelation"][0], Ct, n).numpy()
    want_ar = autocov_feature("ar_coefficient", feats["ar_coefficient"][0], Ct, n).numpy()
    np.testing.assert_array_equal(f[:, 0], want_agg[:, 0].astype(np.float32))
    np.testing.assert_array_equal(f[:, 1], want_ar[:, 0].astype(np.float32))
    adf = np.array([O.comprehensive(x[:, b, 0], "augmented_dickey_fuller", {"attr": "teststat"})
                    for b in range(x.shape[1])])
    np.testing.assert_allclose(f[:, 2], adf.astype(np.float32), rtol=2e-6, atol=1e-6)


def _alt_degenerate(xs, p):
    """Whether the chunk aggregates agg_linear_trend fits are constant to
    rounding (a ramp's chunk variances): its rvalue is then a ratio of two
    rounding residues on either side."""
    cs = p["chunk_size"]
    nc = xs.size // cs
    if nc < 2:
        return False
    ch = xs[: nc * cs].reshape(nc, cs)
    a = {"max": ch.max(1), "min": ch.min(1), "mean": ch.mean(1), "var": ch.var(1)}[p["f_agg"]]
    return np.ptp(a) <= 1e-9 * (np.abs(a).max() + 1e-300)


def _lib_tr_benford():
    from pybasin_amd.window_features import TR

    return TR["BENFORD"]


def _check_trends(out, x, tag, nseries=None):
    """bh_window_trends' features against the oracle (the cubic fit: the
    kernel's elimination against the oracle's LAPACK solve, 1e-7)."""
    from pybasin_amd.window_features import TRENDS_FEATURES, trends_feature

    N, B, S = x.shape
    bs = range(B if nseries is None else min(B, nseries))
    for name, grid in TRENDS_FEATURES.items():
        for p in (grid or [None]):
            got = trends_feature(name, p, out).numpy()
            for b in bs:
                for s in range(S):
                    want = O.comprehensive(x[:, b, s], name, p)
                    g = float(got[b, s])
                    if not math.isfinite(want):
                        assert not math.isfinite(g), (tag, name, p, b, s)
                        continue
                    tol = 1e-7 if name == "friedrich_coefficients" else 1e-11
                    if name == "agg_linear_trend" and p["attr"] == "stderr":
                        tol = 1e-7  # a perfect fit's residual is rounding, under a square root
                    if name == "agg_linear_trend" and p["attr"] == "rvalue" and \
                            _alt_degenerate(x[:, b, s], p):
                        continue
                    assert g == pytest.approx(want, rel=1e-7, abs=tol), (tag, name, p, b, s, g, want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_trends_kernel_equals_oracle_synthetic(cuda_device, dtype):
    """agg_linear_trend (chunk merges of 5 -> 10 -> 50, tails, fewer than two
    chunks), the cubic fit and its sign changes, Benford digits, wavelet
    coefficients (reflect and replicate padding, coefficients past short
    windows) against the oracle; the CPU path agrees."""
    from pybasin_amd.window_features import window_trends

    rng = np.random.default_rng(41)
    cases = list(_synthetic_windows()) + [
        ("walk", np.cumsum(rng.normal(size=(523, 6, 2)), 0) * 0.3),
        ("digits", rng.lognormal(0, 3, size=(200, 5, 1)) * rng.choice([-1.0, 1.0], size=(200, 5, 1))),
    ]
    # leading digits of powers of ten and their neighbours, values outside
    # [1e-30, 1e30] (the fallback of the fast digit path): Benford only (the
    # cubic fit of such a window is meaningless)
    dec = np.array([1e-35, 3e33, 1e6, 999999.9999999999, 0.3, 0.1, 1e-3, 2e-30, 9.5e29, 7.0, 10.0,
                    99.99999999999999, 1e30, 5e-31, 0.0, -4e12])
    dwin = np.stack([np.tile(dec, 4), np.tile(dec[::-1], 4)], 1)[:, :, None]
    for dt_ in (torch.float64,) if dtype == torch.float64 else (torch.float32,):
        dw = torch.tensor(dwin, dtype=dt_)
        got = window_trends(dw.cuda(), 0, False).cpu()[_lib_tr_benford()].numpy()
        for b in range(2):
            want = O.comprehensive(dw[:, b, 0].double().numpy(), "benford_correlation")
            assert got[b, 0] == pytest.approx(want, rel=1e-12, abs=1e-14), (b, got[b, 0], want)
    for tag, w in cases:
        wt = torch.tensor(w, dtype=dtype)
        x = wt.double().numpy()
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            out = window_trends(wt.cuda(), start, True).cpu()
            _check_trends(out, x[start:], f"{tag}/{start}")
            short = window_trends(wt.cuda(), start, False).cpu()
            assert short.shape[0] == 54
            assert torch.equal(torch.nan_to_num(short, 7.0), torch.nan_to_num(out[:54], 7.0)), tag


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_trends_on_integrated_duffing(cuda_device, dtype):
    """The chunk-trend / fit / digit / wavelet features of integrated Duffing
    windows against the oracle (16 series) and the CPU path (all); the
    extractor's columns equal the kernel's."""
    from pybasin_amd.window_features import trends_feature, window_trends

    res = _duffing(dtype)
    out = window_trends(res.y, res.steady_idx, True).cpu()
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    _check_trends(out, x, "duffing", nseries=8)
    cpu = window_trends(res.y.cpu(), res.steady_idx, True)
    np.testing.assert_allclose(out[:48].numpy(), cpu[:48].numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out[53:].numpy(), cpu[53:].numpy(), rtol=1e-9, atol=1e-12)
    feats = {"agg_linear_trend": [{"attr": "slope", "chunk_size": 50, "f_agg": "var"}],
             "cwt_coefficients": [{"widths": (5,), "coeff": 7, "w": 5}],
             "benford_correlation": None}
    fx = StatFeatureExtractor(time_steady=300.0, features=feats)
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    want = np.stack([trends_feature(nm, p, out).numpy()[:, s] for s in range(2)
                     for nm, ps in feats.items() for p in (ps or [None])], 1)
    np.testing.assert_array_equal(f, want.astype(np.float32))


def _check_peaks(counts, perm, x, tag):
    from pybasin_amd.window_features import PEAK_FEATURES, peak_feature

    N, B, S = x.shape
    plist = [("number_peaks", {"n": k}) for k in (10, 50, 1, 64)] + \
        [(nm, p) for nm, g in PEAK_FEATURES.items() for p in g] + \
        [("number_cwt_peaks", {"max_width": 8})]
    for name, p in plist:
        got = peak_feature(name, p, counts, perm, N).numpy()
        for b in range(B):
            for s in range(S):
                want = O.comprehensive(x[:, b, s], name, p)
                g = float(got[b, s])
                if not math.isfinite(want):
                    assert not math.isfinite(g), (tag, name, p, b, s)
                else:
                    assert g == pytest.approx(want, rel=1e-12, abs=1e-12), (tag, name, p, b, s, g, want)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_peaks_and_patterns_kernels_equal_oracle(cuda_device, dtype):
    """Strict peaks of supports 1, 10, 50, 64 (runs capped at 64), plateau
    maxima of widths 1..8, permutation entropies of dimensions 3..7 (ties in
    index order, the reference's colliding pattern ids) against the oracle;
    the CPU path agrees exactly."""
    from pybasin_amd.window_features import window_peak_counts, window_perm_entropy

    rng = np.random.default_rng(43)
    cases = list(_synthetic_windows()) + [
        ("smooth", np.sin(np.arange(700)[:, None, None] * rng.uniform(0.01, 0.3, size=(1, 6, 2)))),
        ("plateaus", np.round(np.sin(np.arange(400)[:, None, None] * 0.05 + rng.normal(size=(1, 5, 2))) * 3))]
    for tag, w in cases:
        wt = torch.tensor(w, dtype=dtype)
        x = wt.double().numpy()
        for start in (0, 1) if w.shape[0] > 1 else (0,):
            c = window_peak_counts(wt.cuda(), start, [10, 50, 1, 64]).cpu()
            counts = {10: c[0], 50: c[1], 1: c[2], 64: c[3], "widths": c[4:]}
            perm = window_perm_entropy(wt.cuda(), start).cpu()
            _check_peaks(counts, perm, x[start:], f"{tag}/{start}")
            cc = window_peak_counts(wt, start, [10, 50, 1, 64])
            assert torch.equal(torch.nan_to_num(c, 7.0), torch.nan_to_num(cc, 7.0)), tag
            cp = window_perm_entropy(wt, start)
            np.testing.assert_allclose(np.nan_to_num(perm.numpy(), nan=7.0),
                                       np.nan_to_num(cp.numpy(), nan=7.0), rtol=1e-12, atol=1e-12)


def test_window_perm_entropy_limit_is_an_error(cuda_device):
    from pybasin_amd.window_features import window_perm_entropy

    with pytest.raises(ValueError, match="LDS budget"):
        window_perm_entropy(torch.zeros(4053, 3, 1, dtype=torch.float32, device="cuda"), 0)
    window_perm_entropy(torch.zeros(4052, 3, 1, dtype=torch.float32, device="cuda"), 0)


@pytest.mark.parametrize("dtype,rows", [(torch.float32, 5290), (torch.float64, 3968)])
def test_window_peaks_at_the_lds_limit(cuda_device, dtype, rows):
    """Windows at the longest length the peak-count kernel stages (16-bit
    candidate positions past 4,096 rows) equal the CPU path exactly and the
    oracle's strict peak counts; one row more is a ValueError."""
    from pybasin_amd.window_features import window_peak_counts

    rng = np.random.default_rng(53)
    t = np.arange(rows)[:, None, None]
    w = np.sin(t * rng.uniform(0.01, 0.4, size=(1, 3, 2))) + 0.3 * rng.normal(size=(rows, 3, 2))
    wt = torch.tensor(w, dtype=dtype)
    c = window_peak_counts(wt.cuda(), 0, [10, 50, 1, 64]).cpu()
    cc = window_peak_counts(wt, 0, [10, 50, 1, 64])
    assert torch.equal(c, cc)
    x = wt.double().numpy()
    for b, s in ((0, 0), (2, 1)):
        for k, sup in enumerate((10, 50, 1, 64)):
            assert float(c[k, b, s]) == O.comprehensive(x[:, b, s], "number_peaks", {"n": sup}), (sup, b, s)
    with pytest.raises(ValueError, match=f"at most {rows} rows"):
        window_peak_counts(torch.zeros(rows + 1, 3, 1, dtype=dtype, device="cuda"), 0, [1, 1, 1, 1])


def test_window_peaks_on_integrated_duffing(cuda_device):
    """Wide-support peaks, cwt peaks and permutation entropies of the
    integrated Duffing windows through the extractor equal the CPU path."""
    res = _duffing(torch.float32)
    feats = {"number_peaks": [{"n": 10}, {"n": 50}, {"n": 3}], "number_cwt_peaks": [{"max_width": 5}],
             "permutation_entropy": [{"tau": 1, "dimension": 5}]}
    fx = StatFeatureExtractor(time_steady=300.0, features=feats)
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    sol_cpu = Solution(res.y[0].cpu(), res.t.cpu(), res.y.cpu())
    g = StatFeatureExtractor(time_steady=300.0, features=feats).extract_features(sol_cpu).numpy()
    np.testing.assert_array_equal(f, g)
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    for b in range(4):
        assert f[b, 0] == O.comprehensive(x[:, b, 0], "number_peaks", {"n": 10})
        assert f[b, 4] == pytest.approx(O.comprehensive(x[:, b, 0], "permutation_entropy",
                                                        {"tau": 1, "dimension": 5}), rel=1e-6)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_embedding_kernels_equal_oracle(cuda_device, dtype):
    """correlation_dimension and the Rosenstein slope (lyapunov_r) of noisy
    sinusoids, noise, the logistic map, constants, short and non-finite
    windows against the oracle; the CPU path agrees."""
    from pybasin_amd.window_features import window_embedding

    rng = np.random.default_rng(47)
    n = 260
    logistic = np.empty(n)
    logistic[0] = 0.3
    for i in range(1, n):
        logistic[i] = 4 * logistic[i - 1] * (1 - logistic[i - 1])
    w = np.concatenate([np.sin(np.arange(n)[:, None, None] * rng.uniform(0.05, 0.5, size=(1, 6, 2)))
                        + 0.05 * rng.normal(size=(n, 6, 2)), rng.normal(size=(n, 2, 2)),
                        np.stack([logistic, logistic[::-1]], 1)[:, None, :],
                        np.full((n, 1, 2), 0.5)], 1)
    w[17, 9, 1] = np.nan
    cases = [("mixed", w), ("short", rng.normal(size=(12, 3, 2))), ("tiny", rng.normal(size=(3, 2, 1)))]
    for tag, win in cases:
        wt = torch.tensor(win, dtype=dtype)
        x = wt.double().numpy()
        for which, name in (("corr_dim", "correlation_dimension"), ("lyap_r", "lyapunov_r")):
            got = window_embedding(wt.cuda(), 0, which).cpu().numpy()
            B, S = got.shape
            for b in range(B):
                for s in range(S):
                    want = O.comprehensive(x[:, b, s], name, {"tau": 1.0})
                    g = float(got[b, s])
                    if not math.isfinite(want):
                        assert not math.isfinite(g), (tag, name, b, s, g)
                    else:
                        assert g == pytest.approx(want, rel=1e-9, abs=1e-12), (tag, name, b, s, g, want)
            cpu = window_embedding(wt, 0, which).numpy()
            np.testing.assert_allclose(np.nan_to_num(got, nan=7.0), np.nan_to_num(cpu, nan=7.0),
                                       rtol=1e-9, atol=1e-12)


def test_window_embedding_on_integrated_duffing(cuda_device):
    """Both estimates on integrated Duffing windows (kept float32 trajectory)
    against the oracle on six series, and through the extractor."""
    from pybasin_amd.window_features import window_embedding

    res = _duffing(torch.float32, n=300)
    x = res.y.double().cpu().numpy()[res.steady_idx:]
    cd = window_embedding(res.y, res.steady_idx, "corr_dim").cpu().numpy()
    ly = window_embedding(res.y, res.steady_idx, "lyap_r").cpu().numpy()
    for b in range(3):
        for s in range(2):
            assert cd[b, s] == pytest.approx(O.comprehensive(x[:, b, s], "correlation_dimension", {}),
                                             rel=1e-9)
            assert ly[b, s] == pytest.approx(O.comprehensive(x[:, b, s], "lyapunov_r", {"tau": 1.0}),
                                             rel=1e-9, abs=1e-12)
    fx = StatFeatureExtractor(time_steady=300.0, features={
        "lyapunov_r": [{"emb_dim": 10, "lag": 1, "trajectory_len": 20, "tau": 2.0}]})
    f = fx.extract_features(Solution(res.y[0], res.t, res.y)).cpu().numpy()
    np.testing.assert_array_equal(f[:, 0], (ly[:, 0] / 2.0).astype(np.float32))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_lempel_ziv_kernel_equals_oracle(cuda_device, dtype):
    """Phrase counts for bins 2, 3, 5, 10, 100 (and eight settings at once) on
    the synthetic windows, a 1,024-row window and integrated Duffing windows
    equal the oracle (the reference's set-of-substrings parse) exactly."""
    from pybasin_amd.window_features import window_lempel_ziv

    rng = np.random.default_rng(53)
    res = _duffing(dtype, n=64)
    duff = res.y.cpu().double().numpy()[res.steady_idx:]
    cases = list(_synthetic_windows()) + [("big", rng.normal(size=(1024, 3, 2))), ("duffing", duff)]
    for tag, w in cases:
        wt = torch.tensor(w, dtype=dtype)
        sd = np.float32 if dtype == torch.float32 else np.float64
        for bins in ([2, 3, 5, 10, 100], [1, 2, 4, 7, 16, 33, 64, 255]):
            if tag == "big" and len(bins) == 8:
                continue  # eight tables of a 1,024-row window exceed the LDS budget
            got = window_lempel_ziv(wt.cuda(), 0, bins).cpu().numpy()
            x = wt.numpy()
            for k, nb in enumerate(bins):
                for b in range(min(w.shape[1], 16)):
                    for s in range(w.shape[2]):
                        want = O.comprehensive(x[:, b, s], "lempel_ziv_complexity", {"bins": nb}, sd)
                        g = float(got[k, b, s])
                        if not math.isfinite(want):
                            assert not math.isfinite(g), (tag, nb, b, s)
                        else:
                            assert g == pytest.approx(want, rel=1e-12), (tag, nb, b, s, g, want)
    with pytest.raises(ValueError, match="1,024 rows"):
        window_lempel_ziv(torch.zeros(1025, 2, 1, dtype=dtype, device="cuda"), 0, [2])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_window_lyapunov_e_kernel_equals_oracle(cuda_device, dtype):
    """Eckmann's four exponents (neighbours within the 8th smallest Chebyshev
    distance, least-squares companion rows, the Householder QR chain with
    positive diagonals) of noisy sinusoids, noise, the logistic map, short,
    constant and non-finite windows against the oracle; the CPU path agrees;
    the extractor's four columns equal the kernel's over tau."""
    from pybasin_amd.window_features import window_embedding

    rng = np.random.default_rng(59)
    n = 220
    logistic = np.empty(n)
    logistic[0] = 0.3
    for i in range(1, n):
        logistic[i] = 4 * logistic[i - 1] * (1 - logistic[i - 1])
    w = np.concatenate([np.sin(np.arange(n)[:, None, None] * rng.uniform(0.05, 0.5, size=(1, 4, 2)))
                        + 0.05 * rng.normal(size=(n, 4, 2)), rng.normal(size=(n, 2, 2)),
                        np.stack([logistic, logistic[::-1]], 1)[:, None, :],
                        np.full((n, 1, 2), 0.5)], 1)
    w[30, 5, 0] = np.inf
    for tag, win in (("mixed", w), ("short", rng.normal(size=(20, 3, 2))), ("tiny", rng.normal(size=(12, 2, 1)))):
        wt = torch.tensor(win, dtype=dtype)
        x = wt.double().numpy()
        got = window_embedding(wt.cuda(), 0, "lyap_e").cpu().numpy()
        for b in range(win.shape[1]):
            for s in range(win.shape[2]):
                want = np.array(O.lyapunov_e(x[:, b, s]))
                g = got[:, b, s]
                fin = np.isfinite(want)
                assert np.array_equal(np.isfinite(g), fin), (tag, b, s, g, want)
                np.testing.assert_allclose(g[fin], want[fin], rtol=1e-7, atol=1e-9, err_msg=f"{tag} {b} {s}")
        cpu = window_embedding(wt, 0, "lyap_e").numpy()
        np.testing.assert_allclose(np.nan_to_num(got, posinf=7.0), np.nan_to_num(cpu, posinf=7.0),
                                   rtol=1e-7, atol=1e-9)
    t = torch.linspace(0, 10, n, dtype=torch.float64)
    y = torch.tensor(w[:, :3], dtype=dtype).cuda()
    fx = StatFeatureExtractor(time_steady=0.0, features={
        "lyapunov_e": [{"emb_dim": 10, "matrix_dim": 4, "min_nb": 8, "min_tsep": 0, "tau": 0.5}]})
    f = fx.extract_features(Solution(y[0], t.cuda(), y)).cpu().numpy()
    assert fx.feature_names[3] == "state_0__lyapunov_e__emb_dim_10__matrix_dim_4__min_nb_8__min_tsep_0__tau_0.5__3"
    ex = window_embedding(y, 0, "lyap_e").cpu().numpy() / 0.5
    np.testing.assert_array_equal(f[:, :4], ex[:, :, 0].T.astype(np.float32))


# the reference's own peak cases (tests/unit/ts_torch/calculators/
# test_torch_features_pattern.py:18-58, 207-262): series and their strict
# peak counts of support 1
_REFERENCE_PEAK_CASES = [
    ([0.0, 1.0, 0.0, 2.0, 0.0], 2),
    ([0.0, 1.0, 2.0, 3.0, 4.0, 10.0, 4.0, 3.0, 2.0, 1.0], 1),
    (list(range(10)), 0),
    ([5.0] * 10, 0),
    ([0.0, 1.0, 0.0], 1),
    ([0.0, 1.0, 0.5, 2.0, 0.0], 2),
    ([4.0, 3.0, 2.0, 1.0, 0.0], 0),
    ([0.0, 2.0, 1.0, 2.0, 0.0], 2),
    ([0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0], 3),
    ([0.0, 1.0, 2.0, 2.0, 2.0, 1.0, 0.0], 0),
]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_reference_peak_cases(cuda_device, dtype):
    """number_peaks (n = 1) through the window-feature slots and through the
    peak-count kernel equals the reference tests' counts; a sinusoid's
    count equals scipy.signal.argrelmax's (the reference's own check)."""
    from scipy.signal import argrelmax

    from pybasin_amd.window_features import comprehensive_feature, window_feature_block, window_peak_counts

    t = np.linspace(0, 4 * np.pi, 1001)
    cases = _REFERENCE_PEAK_CASES + [(list(np.sin(t)), len(argrelmax(np.sin(t))[0]))]
    for data, want in cases:
        y = torch.tensor(data, dtype=dtype).view(-1, 1, 1).cuda()
        blk = window_feature_block(y, 0).cpu()
        assert comprehensive_feature("number_peaks", {"n": 1}, blk, y.shape[0])[0, 0] == want, data
        assert window_peak_counts(y, 0, [1])[0, 0, 0].item() == want, data
