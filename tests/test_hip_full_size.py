"""Full-size workloads on the GPU, checked through size-independent properties.

IC-by-IC comparisons with the oracle at these sizes are in
test_hip_parity_full.py; these tests check what must hold at any size: each IC's result does not
depend on where it sits in the batch or on the launch shape (bitwise), the
float32 and float64 runs agree on the labels away from basin boundaries, and
the basin-stability fractions agree with the reference's own 4,000-sample
runs (tests/integration/duffing/main_duffing_supervised.json,
tests/integration/lorenz/main_lorenz.json) within their combined binomial
errors.  Launch shapes include the two-phase launch.  Small-size edge cases (empty batch, one IC, ragged counts with
refills) close the file.
"""

import json
import math
import os

import numpy as np
import pytest
import torch

from pybasin_amd import _lib
from pybasin_amd.cluster import knn
from pybasin_amd.solvers import HipDopri5Solver
from pybasin_amd.systems import DuffingODE, LorenzODE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")

DUFF = {"delta": 0.08, "k3": 1.0, "A": 0.2}
LOR = {"sigma": 0.12, "r": 0.0, "b": -0.6}
DUFF_TEMPLATES = [[-0.21, 0.02], [1.05, 0.77], [-0.67, 0.02], [-0.46, 0.30], [-0.43, 0.12]]
LOR_TEMPLATES = [[0.8, -3.0, 0.0], [-0.8, 3.0, 0.0]]


def _fixture_bs(key):
    with open(os.path.join(GOLD, "label_fixtures.json")) as f:
        fx = json.load(f)[key]
    return [(b["basinStability"], b["standardError"]) for b in fx["bs"]], fx["n"]


def _uniform(n, lo, hi, seed):
    g = torch.Generator().manual_seed(seed)
    lo, hi = torch.tensor(lo), torch.tensor(hi)
    return torch.rand(n, len(lo), generator=g) * (hi - lo) + lo


def _duffing(y0, dtype, **kw):
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             cache_dir=None, **kw)
    batch = torch.cat([y0, torch.tensor(DUFF_TEMPLATES)]).to(dtype).cuda()
    save_ts, batch = solver._prepare_tensors(batch)
    return solver.run_fused(DuffingODE(DUFF), batch, save_ts, None, keep_trajectories=False,
                            steady_idx=0)


def _duffing_labels(res, n):
    feats = torch.stack([res.stats[:, 0, 3], res.stats[:, 0, 1]], 1).contiguous()
    return knn(feats[:n].contiguous(), feats[n:].contiguous(), 1)[1][:, 0]


def _z(bs_got, n_got, ref, n_ref):
    """Largest |difference| / combined binomial standard error over labels."""
    worst = 0.0
    for got, (want, _) in zip(bs_got, ref):
        se = math.sqrt(want * (1 - want) / n_ref + max(got * (1 - got), 1e-12) / n_got)
        worst = max(worst, abs(got - want) / se)
    return worst


def test_duffing_100k_permutation_invariance_and_reference_fractions(cuda_device):
    n = 100_000
    y0 = _uniform(n, [-1.0, -0.5], [1.0, 1.0], 0)
    base = _duffing(y0, torch.float32)
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(1))
    shuffled = _duffing(y0[perm], torch.float32)
    inv = torch.argsort(perm).cuda()
    assert torch.equal(shuffled.stats[:n][inv], base.stats[:n])
    assert torch.equal(shuffled.nsteps[:n][inv], base.nsteps[:n])
    # the two-phase launch (long trajectories packed after 2,800 steps)
    phased = _duffing(y0, torch.float32, phase_iters=2800)
    assert torch.equal(phased.stats, base.stats) and torch.equal(phased.nsteps, base.nsteps)
    assert (base.status == _lib.BH_LANE_OK).all()
    lab = _duffing_labels(base, n)
    bs = (torch.bincount(lab, minlength=5).double() / n).cpu().tolist()
    ref, n_ref = _fixture_bs("duffing_main")
    z = _z(bs, n, ref, n_ref)
    print(f"[result] duffing 100k BS {np.round(bs, 5).tolist()} ref {[r[0] for r in ref]} worst z {z:.2f}",
          flush=True)
    assert z < 3.0


def test_duffing_100k_float32_and_float64_labels_agree(cuda_device):
    n = 100_000
    y0 = _uniform(n, [-1.0, -0.5], [1.0, 1.0], 0)
    l32 = _duffing_labels(_duffing(y0, torch.float32), n)
    r64 = _duffing(y0, torch.float64)
    assert (r64.status == _lib.BH_LANE_OK).all()
    l64 = _duffing_labels(r64, n)
    agree = float((l32 == l64).double().mean())
    print(f"[result] duffing 100k float32 vs float64 label agreement {agree:.5f}", flush=True)
    # ICs whose label flips with the working precision sit on basin
    # boundaries of this multistable system; they are a small minority
    assert agree >= 0.995


def test_lorenz_200k_launch_shapes_agree_and_reference_fractions(cuda_device):
    n = 200_000
    y0 = _uniform(n, [-10.0, -20.0, 0.0], [10.0, 20.0, 0.0], 0)
    batch = torch.cat([y0, torch.tensor(LOR_TEMPLATES)]).cuda()
    out = []
    for cap, q, k in ((0, 0, 0), (131072, 1, 0), (65536, 64, 0), (0, 0, 600)):
        solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                                 cache_dir=None, stop_threshold=200.0, inf_fill=True, max_lanes=cap,
                                 refill_quorum=q, phase_iters=k)
        save_ts, b = solver._prepare_tensors(batch)
        out.append(solver.run_fused(LorenzODE(LOR), b, save_ts, None, keep_trajectories=False,
                                    steady_idx=0))
    for r in out[1:]:
        assert torch.equal(r.stats.nan_to_num(7.0), out[0].stats.nan_to_num(7.0))
        assert torch.equal(r.status, out[0].status) and torch.equal(r.nsteps, out[0].nsteps)
    res = out[0]
    mean_x = res.stats[:, 0, 0:1].contiguous()
    lab = knn(mean_x[:n], mean_x[n:], 1)[1][:, 0]
    lab = torch.where(res.status[:n] == _lib.BH_LANE_DIVERGED, torch.full_like(lab, 2), lab)
    bs = (torch.bincount(lab, minlength=3).double() / n).cpu().tolist()
    ref, n_ref = _fixture_bs("lorenz_main")
    z = _z(bs, n, ref[:3], n_ref)
    print(f"[result] lorenz 200k BS {np.round(bs, 5).tolist()} ref {[r[0] for r in ref[:3]]} worst z {z:.2f}",
          flush=True)
    assert z < 3.0


def test_empty_single_and_ragged_batches(cuda_device):
    y0 = _uniform(257, [-1.0, -0.5], [1.0, 1.0], 3)
    s = HipDopri5Solver(t_span=(0, 50), t_steps=50, device="cuda", cache_dir=None)
    empty = s.integrate_fused(DuffingODE(DUFF), y0[:0].cuda())
    assert empty.stats.shape[0] == 0 and empty.status.shape == (0,)
    full = s.integrate_fused(DuffingODE(DUFF), y0.cuda(), keep_trajectories=True)
    one = s.integrate_fused(DuffingODE(DUFF), y0[100:101].cuda(), keep_trajectories=True)
    assert torch.equal(one.y[:, 0], full.y[:, 100]) and torch.equal(one.stats[0], full.stats[100])
    for cap, q in ((64, 1), (64, 64), (192, 5)):
        r = HipDopri5Solver(t_span=(0, 50), t_steps=50, device="cuda", cache_dir=None, max_lanes=cap,
                            refill_quorum=q).integrate_fused(DuffingODE(DUFF), y0.cuda(),
                                                             keep_trajectories=True)
        assert torch.equal(r.y, full.y) and torch.equal(r.nsteps, full.nsteps)
