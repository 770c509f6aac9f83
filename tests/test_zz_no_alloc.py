"""bh_solve_batch_phased_noalloc (scratch in a static device buffer) against
bh_solve_batch_phased (caller workspace) through the C-ABI: the same
kernels, so the outputs are equal bit for bit, on two calls in a row, the
second with a smaller input (the shared buffer keeps the first call's
contents)."""

import numpy as np
import pytest
import torch

from pybasin_amd import _lib

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    g = torch.Generator().manual_seed(seed)
    lo, hi = torch.tensor([-1.0, -0.5]), torch.tensor([1.0, 1.0])
    return (torch.rand(n, 2, generator=g) * (hi - lo) + lo).cuda()


def _call(noalloc, y0, save, params, phase_iters):
    n = y0.shape[0]
    out = dict(feats=torch.empty((n, 2, _lib.BH_NFEAT), dtype=torch.float64, device="cuda"),
               final=torch.empty((n, 2), dtype=torch.float32, device="cuda"),
               status=torch.empty((n,), dtype=torch.int32, device="cuda"),
               nsteps=torch.empty((n,), dtype=torch.int64, device="cuda"),
               traj=torch.empty((save.shape[0], n, 2), dtype=torch.float32, device="cuda"))
    L = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    common = (_lib.BH_SYS_DUFFING, _lib.BH_F32, y0.data_ptr(), n, 2, params.data_ptr(), 3, 0, 0.0,
              save.data_ptr(), save.shape[0], 1e-8, 1e-6, 1 << 30, 0, 0.0, save.shape[0] // 2,
              out["traj"].data_ptr(), out["feats"].data_ptr(), out["final"].data_ptr(),
              out["status"].data_ptr(), out["nsteps"].data_ptr(), phase_iters)
    if noalloc:
        rc = L.bh_solve_batch_phased_noalloc(*common, s)
    else:
        need = L.bh_solve_batch_phased_workspace_bytes(_lib.BH_SYS_DUFFING, _lib.BH_F32, 2, n)
        ws = torch.empty((need,), dtype=torch.uint8, device="cuda")
        rc = L.bh_solve_batch_phased(*common, ws.data_ptr(), need, s)
        out["ws"] = ws
    _lib.check(rc, "noalloc test")
    torch.cuda.synchronize()
    return out


def test_noalloc_equals_workspace_entry_on_two_calls(cuda_device):
    save = torch.linspace(150.0, 200.0, 100, dtype=torch.float64, device="cuda")
    params = torch.tensor([[0.08, 1.0, 0.2]], dtype=torch.float64, device="cuda")
    for n, seed, k in ((60_000, 1, 600), (3_001, 2, 300)):
        print(f"[case] noalloc Duffing n={n} phase_iters={k}", flush=True)
        y0 = _inputs(n, seed)
        a = _call(True, y0, save, params, k)
        b = _call(False, y0, save, params, k)
        for key in ("feats", "final", "status", "nsteps", "traj"):
            assert torch.equal(a[key].nan_to_num(7.0), b[key].nan_to_num(7.0)), (n, key)
        assert (a["status"] == 0).all()
        # some trajectories were suspended and resumed (the scratch held them)
        assert int((a["nsteps"] > k).sum()) > 0
