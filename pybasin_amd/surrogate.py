"""Learned surrogate: (theta, omega, T) -> attractor label, with Dopri5 fallback.

The network (3 -> 256 -> 256 -> 256 -> 2, SiLU) is trained and evaluated by
the HIP kernels of libbasinhip.so (bf16 weights and activations, float32
accumulation on the matrix cores, Adam on float32 master weights).  ICs whose
softmax confidence is below a gate are integrated by the per-lane Dopri5
kernel; a random audit sample is integrated regardless and reports the
surrogate's label error and basin-stability error against full integration.

The surrogate has no reference counterpart; its labels come from the
integrator path that replaces TorchDiffEqSolver (the pendulum of
case_studies/pendulum/pendulum_ode.py:18-40 with the 1-NN template
classifier of setup_pendulum_system.py:44-65).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import _lib

HIDDEN = 256
IN_PAD = 16
N_IN = 6  # sin theta, cos theta, omega, T, theta, energy (see bh_mlp_prep_input)


def _s(stream):
    return stream.cuda_stream


class SurrogateMLP:
    """Parameters and kernels of the 4-layer MLP on one device."""

    def __init__(self, device="cuda:0", hidden: int = HIDDEN, seed: int = 0,
                 norm=(0.0, 1.0 / math.pi, 0.0, 0.1, 0.55, 1.0 / 0.45)):
        self.device = torch.device(device)
        self.hidden = hidden
        self.norm = tuple(float(v) for v in norm)
        g = torch.Generator().manual_seed(seed)
        dims = [(IN_PAD, hidden, N_IN), (hidden, hidden, hidden), (hidden, hidden, hidden), (hidden, 2, hidden)]
        self.W, self.b = [], []
        for fin, fout, fan_in in dims:
            bound = 1.0 / math.sqrt(fan_in)
            w = (torch.rand(fin, fout, generator=g) * 2 - 1) * bound
            if fin == IN_PAD:
                w[N_IN:] = 0.0  # padded input columns
            b = (torch.rand(fout, generator=g) * 2 - 1) * bound
            self.W.append(w.to(self.device).contiguous())
            self.b.append(b.to(self.device).contiguous())
        self.Wb = [w.to(torch.bfloat16).contiguous() for w in self.W]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step = 0

    # ---- kernels ---------------------------------------------------------
    def _gemm(self, A, sam, sak, B, sbk, sbn, C, M, N, K, stream, accumulate=0):
        _lib.check(_lib.lib().bh_gemm_bf16(A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn,
                                           C.data_ptr(), C.stride(0), M, N, K, accumulate,
                                           _s(stream)), "bh_gemm_bf16")

    def prep(self, x: torch.Tensor, stream) -> torch.Tensor:
        M = x.shape[0]
        out = torch.empty((M, IN_PAD), dtype=torch.bfloat16, device=self.device)
        c0, s0, c1, s1, c2, s2 = self.norm
        _lib.check(_lib.lib().bh_mlp_prep_input(x.data_ptr(), M, c0, s0, c1, s1, c2, s2,
                                                out.data_ptr(), _s(stream)), "bh_mlp_prep_input")
        return out

    def forward(self, x: torch.Tensor, stream=None, keep: bool = False):
        """x (M, 3) float32 on device -> logits (M, 2) float32 (bias added).
        keep=True also returns the saved tensors for backward."""
        stream = stream or torch.cuda.current_stream(self.device)
        x = x.to(self.device, torch.float32).contiguous()
        M = x.shape[0]
        a = self.prep(x, stream)
        acts, zs = [a], []
        for layer in range(3):
            W = self.Wb[layer]
            K, N = W.shape
            Z = torch.empty((M, N), dtype=torch.float32, device=self.device)
            self._gemm(acts[-1], acts[-1].stride(0), 1, W, N, 1, Z, M, N, K, stream)
            zsave = torch.empty_like(Z) if keep else None
            A = torch.empty((M, N), dtype=torch.bfloat16, device=self.device)
            _lib.check(_lib.lib().bh_mlp_bias_act(Z.data_ptr(), M, N, self.b[layer].data_ptr(), 1,
                                                  _lib.ptr(zsave), A.data_ptr(), N, None,
                                                  _s(stream)), "bh_mlp_bias_act")
            zs.append(zsave)
            acts.append(A)
        Z4 = torch.empty((M, 2), dtype=torch.float32, device=self.device)
        self._gemm(acts[-1], self.hidden, 1, self.Wb[3], 2, 1, Z4, M, 2, self.hidden, stream)
        logits = torch.empty_like(Z4)
        _lib.check(_lib.lib().bh_mlp_bias_act(Z4.data_ptr(), M, 2, self.b[3].data_ptr(), 0, None,
                                              None, 0, logits.data_ptr(), _s(stream)), "bh_mlp_bias_act")
        if keep:
            return logits, (acts, zs)
        return logits

    def gradients(self, x: torch.Tensor, labels: torch.Tensor, stream=None):
        """Mean cross-entropy loss and its gradients (dW list, db list)."""
        stream = stream or torch.cuda.current_stream(self.device)
        labels = labels.to(self.device, torch.int32).contiguous()
        logits, (acts, zs) = self.forward(x, stream, keep=True)
        M = logits.shape[0]
        L = _lib.lib()
        dlog = torch.empty_like(logits)
        dlogb = torch.empty((M, 2), dtype=torch.bfloat16, device=self.device)
        loss = torch.empty((M,), dtype=torch.float32, device=self.device)
        _lib.check(L.bh_mlp_softmax_ce(logits.data_ptr(), labels.data_ptr(), M, dlog.data_ptr(),
                                       dlogb.data_ptr(), loss.data_ptr(), _s(stream)), "bh_mlp_softmax_ce")
        dW = [None] * 4
        db = [None] * 4
        gz, gzb = dlog, dlogb  # gradient w.r.t. this layer's pre-activation
        for layer in range(3, -1, -1):
            Ain = acts[layer]
            K_in, N_out = self.Wb[layer].shape
            # dW = Ain^T . gz  (K_in x N_out, reduction over M)
            gw = torch.empty((K_in, N_out), dtype=torch.float32, device=self.device)
            self._gemm(Ain, 1, Ain.stride(0), gzb, N_out, 1, gw, K_in, N_out, M, stream)
            gb = torch.empty((N_out,), dtype=torch.float32, device=self.device)
            _lib.check(L.bh_mlp_colsum(gz.data_ptr(), M, N_out, gb.data_ptr(), _s(stream)), "bh_mlp_colsum")
            dW[layer], db[layer] = gw, gb
            if layer == 0:
                break
            # dA_in = gz . W^T  (M x K_in)
            dA = torch.empty((M, K_in), dtype=torch.float32, device=self.device)
            self._gemm(gzb, N_out, 1, self.Wb[layer], 1, N_out, dA, M, K_in, N_out, stream)
            nz = torch.empty_like(dA)
            nzb = torch.empty((M, K_in), dtype=torch.bfloat16, device=self.device)
            _lib.check(L.bh_mlp_silu_bwd(dA.data_ptr(), zs[layer - 1].data_ptr(), M * K_in,
                                         nz.data_ptr(), nzb.data_ptr(), _s(stream)), "bh_mlp_silu_bwd")
            gz, gzb = nz, nzb
        return loss, dW, db

    def adam_step(self, dW, db, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        stream = stream or torch.cuda.current_stream(self.device)
        self.step += 1
        L = _lib.lib()
        for i in range(4):
            _lib.check(L.bh_mlp_adam(self.W[i].data_ptr(), dW[i].data_ptr(), self.mW[i].data_ptr(),
                                     self.vW[i].data_ptr(), self.W[i].numel(), lr, b1, b2, eps,
                                     self.step, self.Wb[i].data_ptr(), _s(stream)), "bh_mlp_adam")
            _lib.check(L.bh_mlp_adam(self.b[i].data_ptr(), db[i].data_ptr(), self.mb[i].data_ptr(),
                                     self.vb[i].data_ptr(), self.b[i].numel(), lr, b1, b2, eps,
                                     self.step, None, _s(stream)), "bh_mlp_adam")

    def train_step(self, x: torch.Tensor, labels: torch.Tensor, lr=1e-3, b1=0.9, b2=0.999,
                   eps=1e-8, stream=None) -> torch.Tensor:
        """One native training step (bh_mlp_train_step: forward, loss, backward
        and Adam in one call, the same kernels as gradients() + adam_step());
        returns the per-row loss.  The workspace is kept between calls, so
        calls on different streams must not overlap."""
        import ctypes

        stream = stream or torch.cuda.current_stream(self.device)
        x = x.to(self.device, torch.float32).contiguous()
        labels = labels.to(self.device, torch.int32).contiguous()
        M = x.shape[0]
        L = _lib.lib()
        need = L.bh_mlp_train_workspace_bytes(M, self.hidden)
        ws = getattr(self, "_train_ws", None)
        if ws is None or ws.numel() < need:
            ws = self._train_ws = torch.empty((max(need, 1),), dtype=torch.uint8, device=self.device)
        loss = torch.empty((M,), dtype=torch.float32, device=self.device)
        arr = lambda ts: (ctypes.c_void_p * 4)(*[t.data_ptr() for t in ts])  # noqa: E731
        norm = (ctypes.c_float * 6)(*self.norm)
        self.step += 1
        _lib.check(L.bh_mlp_train_step(x.data_ptr(), labels.data_ptr(), M, norm, self.hidden,
                                       arr(self.W), arr(self.Wb), arr(self.b), arr(self.mW),
                                       arr(self.vW), arr(self.mb), arr(self.vb), lr, b1, b2, eps,
                                       self.step, loss.data_ptr(), ws.data_ptr(), ws.numel(),
                                       _s(stream)), "bh_mlp_train_step")
        return loss

    def fit(self, x: torch.Tensor, labels: torch.Tensor, epochs: int = 60, batch: int = 1024,
            lr: float = 3e-3, seed: int = 0) -> list[float]:
        """Mini-batch Adam on (x, labels); returns the mean loss per epoch."""
        x = x.to(self.device, torch.float32).contiguous()
        labels = labels.to(self.device, torch.int32).contiguous()
        n = x.shape[0]
        g = torch.Generator(device=self.device).manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, device=self.device, generator=g)
            tot = torch.zeros((), dtype=torch.float64, device=self.device)
            lr_ep = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))
            for i in range(0, n - batch + 1, batch):
                idx = perm[i:i + batch]
                loss = self.train_step(x[idx], labels[idx], lr=lr_ep)
                tot += loss.double().sum()
            hist.append(float(tot) / (n - n % batch))
        return hist

    def packed(self):
        """Weights in the fragment order of the fused inference kernel
        ([kstep][colblock][lane][8] bf16 per layer, 32-deep steps) and the
        biases concatenated."""
        if getattr(self, "_packed_step", None) == self.step:
            return self._packed
        frags = []
        for W in self.Wb:
            K, N = W.shape
            ks, cb = (K + 31) // 32, (N + 15) // 16
            Wp = torch.zeros((ks * 32, cb * 16), dtype=torch.bfloat16, device=self.device)
            Wp[:K, :N] = W
            # element (s, c, lane, e) = Wp[32 s + 8 (lane // 16) + e, 16 c + lane % 16]
            t = Wp.view(ks, 4, 8, cb, 16)            # s, kq(=lane//16), e, c, lr
            t = t.permute(0, 3, 1, 4, 2).contiguous()  # s, c, kq, lr, e
            frags.append(t.view(-1))
        b = torch.cat([self.b[0], self.b[1], self.b[2], self.b[3]]).contiguous()
        self._packed = (torch.cat(frags).contiguous(), b)
        self._packed_step = self.step
        return self._packed

    def predict(self, x: torch.Tensor, stream=None):
        """labels (int32) and confidences (max softmax probability) for x (M, 3),
        through the fused inference kernel."""
        stream = stream or torch.cuda.current_stream(self.device)
        x = x.to(self.device, torch.float32).contiguous()
        M = x.shape[0]
        lab = torch.empty((M,), dtype=torch.int32, device=self.device)
        conf = torch.empty((M,), dtype=torch.float32, device=self.device)
        if M == 0:
            return lab, conf
        w, b = self.packed()
        c0, s0, c1, s1, c2, s2 = self.norm
        _lib.check(_lib.lib().bh_mlp_infer(x.data_ptr(), M, w.data_ptr(), b.data_ptr(), c0, s0, c1,
                                           s1, c2, s2, lab.data_ptr(), conf.data_ptr(),
                                           _s(stream)), "bh_mlp_infer")
        return lab, conf

    def predict_unfused(self, x: torch.Tensor, chunk: int = 1 << 20, stream=None):
        """labels and confidences through the layer-by-layer kernels."""
        stream = stream or torch.cuda.current_stream(self.device)
        x = x.to(self.device, torch.float32).contiguous()
        M = x.shape[0]
        lab = torch.empty((M,), dtype=torch.int32, device=self.device)
        conf = torch.empty((M,), dtype=torch.float32, device=self.device)
        zero_b = torch.zeros(2, dtype=torch.float32, device=self.device)
        for s in range(0, M, chunk):
            e = min(M, s + chunk)
            logits = self.forward(x[s:e], stream)
            _lib.check(_lib.lib().bh_mlp_classify(logits.data_ptr(), zero_b.data_ptr(), e - s,
                                                  lab[s:].data_ptr(), conf[s:].data_ptr(),
                                                  _s(stream)), "bh_mlp_classify")
        return lab, conf


@dataclass
class GatedResult:
    labels: torch.Tensor          # (n,) int32 final labels
    routed: torch.Tensor          # (n,) bool: below the gate or in a distrusted group
    audit_idx: torch.Tensor       # indices of the audit sample
    audit_surrogate: torch.Tensor  # surrogate labels on the audit sample
    audit_integrated: torch.Tensor  # integrator labels on the audit sample
    integrated: torch.Tensor | None = None  # (n,) bool: integrated over the full span
    group_routed: torch.Tensor | None = None  # (n,) bool: routed by the group audit
    evolved: torch.Tensor | None = None      # (n,) bool: labelled after a short integration
    audit_evolved_error: float | None = None  # evolved-label error on routed audit rows
    n_late: int = 0                           # rows integrated after the network passes
    audit_final: torch.Tensor | None = None   # labels the pass gives the audited rows when not audited


def pendulum_labels_from_stats(stats: torch.Tensor) -> torch.Tensor:
    """FP (0) / LC (1) from the window statistics of omega: FP when
    |max - mean| < 0.01 (case_studies/pendulum/pendulum_feature_extractor.py:17-29)."""
    delta = (stats[:, 1, 3] - stats[:, 1, 0]).abs()
    lab = (delta >= 0.01).to(torch.int32)
    return torch.where(torch.isfinite(delta), lab, torch.ones_like(lab))


def integrate_labels(solver, ode, x: torch.Tensor, time_steady: float, stream=None,
                     with_result: bool = False):
    """Plain path for rows (theta, omega, T): per-lane Dopri5 with the row's
    torque as parameter, fused window statistics, threshold labels.
    with_result=True returns (labels, FusedResult)."""
    dev = solver.device
    x = x.to(dev, torch.float32)
    n = x.shape[0]
    if n == 0:
        empty = torch.empty((0,), dtype=torch.int32, device=dev)
        return (empty, None) if with_result else empty
    alpha, K = float(ode.params["alpha"]), float(ode.params["K"])
    params = torch.stack([torch.full((n,), alpha, dtype=torch.float64, device=dev),
                          x[:, 2].double(), torch.full((n,), K, dtype=torch.float64, device=dev)], 1)
    y0 = x[:, :2].contiguous()
    save_ts, y0 = solver._prepare_tensors(y0)
    res = solver.run_fused(ode, y0, save_ts, params.contiguous(), keep_trajectories=False,
                           time_steady=time_steady, stream=stream)
    lab = pendulum_labels_from_stats(res.stats)
    return (lab, res) if with_result else lab


def tune_plain_launch(solver, ode, x: torch.Tensor, time_steady: float,
                      quantiles=(0.1, 0.25, 0.5), reps: int = 2) -> dict:
    """Launch shape for the plain path on rows (theta, omega, T), set on the
    solver: the whole-device refilling launch or a two-phase launch whose
    step budget is the attempted-step count at which the given fractions of
    the rows have ended (bh_autotune_lane on the rows themselves).  Returns
    the shape with the candidates' times."""
    dev = solver.device
    x = x.to(dev, torch.float32)
    n = x.shape[0]
    alpha, K = float(ode.params["alpha"]), float(ode.params["K"])
    params = torch.stack([torch.full((n,), alpha, dtype=torch.float64, device=dev),
                          x[:, 2].double(), torch.full((n,), K, dtype=torch.float64, device=dev)], 1)
    save_ts, y0 = solver._prepare_tensors(x[:, :2].contiguous())
    plan = solver.plan_fused(ode, y0, save_ts, params.contiguous(), time_steady=time_steady)
    _, _, _, timings, res = solver.tune_fused(plan, [0], [64], reps=reps)
    steps = res.nsteps.double()
    qs = torch.quantile(steps[:: max(1, n // 1_000_000)].contiguous(),
                        torch.tensor(list(quantiles), dtype=torch.float64, device=dev))
    top = int(steps.max())
    budgets = sorted({int(v) + 1 for v in qs.cpu().tolist() if 0 < int(v) + 1 < top})
    if budgets:
        timings.update(solver.tune_fused(plan, [], [], budgets, reps=reps)[3])
    best = min(timings, key=timings.get)
    shape = ({"max_lanes": 0, "refill_quorum": 0, "phase_iters": best[1]} if best[0] == "phased"
             else {"max_lanes": best[1], "refill_quorum": best[2], "phase_iters": 0})
    solver.max_lanes, solver.refill_quorum, solver.phase_iters = (
        shape["max_lanes"], shape["refill_quorum"], shape["phase_iters"])
    names = {k: (f"phase{k[1]}" if k[0] == "phased" else f"{k[1]}/q{k[2]}") for k in timings}
    return dict(shape, timings_ms={names[k]: round(v, 3) for k, v in timings.items()})


def evolve_states(solver, ode, x: torch.Tensor, t_probe: float, stream=None) -> torch.Tensor:
    """Rows (theta, omega, T) advanced by Dopri5 from t_span[0] to t_probe
    (same tolerances, per-row torque), as new (theta, omega, T) rows."""
    dev = solver.device
    n = x.shape[0]
    if n == 0:
        return x
    alpha, K = float(ode.params["alpha"]), float(ode.params["K"])
    params = torch.stack([torch.full((n,), alpha, dtype=torch.float64, device=dev),
                          x[:, 2].double(), torch.full((n,), K, dtype=torch.float64, device=dev)], 1)
    save = torch.full((1,), float(t_probe), dtype=torch.float32, device=dev)  # no host copy
    res = solver.run_fused(ode, x[:, :2].contiguous(), save, params.contiguous(),
                           keep_trajectories=False, steady_idx=1, stream=stream)
    return torch.cat([res.final, x[:, 2:3]], 1).contiguous()


def classify_gated(model: SurrogateMLP, solver, ode, x: torch.Tensor, gate: float = 0.95,
                   audit_frac: float = 0.02, group: torch.Tensor | None = None,
                   time_steady: float = 950.0, seed: int = 0,
                   group_error_cap: float | None = 0.01,
                   evolve_t: float | None = None, early_frac: float = 0.0,
                   stage_times: dict | None = None, n_groups: int | None = None) -> GatedResult:
    """Surrogate labels for confident rows, Dopri5 for the rest.

    A random audit sample (audit_frac of every group) is integrated over the
    full span on a second stream while the network runs.  When it ends, a
    group whose labels disagree with these integrations on more than
    group_error_cap of its audited rows (rows that would have been
    integrated anyway excluded) is integrated entirely.

    evolve_t: rows below the gate are first integrated to t = evolve_t only
    and classified again at the state they reached: the label is a property
    of the trajectory, and after the transient most states lie far from the
    basin boundary.  The audit sample is evolved too, and a group whose first
    answers disagree with the answers at the evolved states on more than
    group_error_cap of its confident audited rows is treated as below the
    gate (its confident rows are evolved as well), a decision available long
    before the audit's full integrations end.  Rows still below the gate
    after evolving are integrated over the full span on the main stream,
    while the audit finishes.  Without evolve_t every below-gate row is
    integrated over the full span.

    early_frac: the early_frac * n least confident rows are integrated over
    the full span in the audit's launch, so the rows a short integration
    leaves undecided (near-boundary ICs, mostly among them) do not need a
    full-span launch of their own afterwards.  A scheduling choice only:
    integrated rows always take the integrator's label.

    stage_times: if a dict is given, it receives the main stream's elapsed
    milliseconds at each stage boundary (HIP events; read after the call),
    the audit launch's span (audit_launch_ms) and step counts
    (audit_nsteps), and the network pass's span (infer_launch_ms).

    n_groups: the number of groups (max(group) + 1), if the caller knows it
    (saves reading it back from the device)."""
    dev = solver.device
    marks = []

    def mark(name):
        if stage_times is not None:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(torch.cuda.current_stream(dev))
            marks.append((name, ev))

    mark("start")
    x = x.to(dev, torch.float32).contiguous()
    n = x.shape[0]
    main = torch.cuda.current_stream(dev)
    g = torch.Generator(device=dev).manual_seed(seed)
    u = torch.rand((n,), generator=g, device=dev)
    audit = u < audit_frac
    aidx = torch.nonzero(audit).squeeze(1)
    if group is not None:  # every group audited (a uniform draw almost always is)
        gl = group.long()
        ng = int(n_groups) if n_groups is not None else int(gl.max()) + 1
        have = torch.zeros((ng,), dtype=torch.bool, device=dev)
        have[gl[aidx]] = True
        missing = torch.nonzero(~have).squeeze(1).tolist()
        if missing:
            for gi in missing:
                audit[torch.nonzero(gl == gi)[0, 0]] = True
            aidx = torch.nonzero(audit).squeeze(1)
    side = torch.cuda.Stream(device=dev)
    early = torch.zeros_like(audit)
    k = int(early_frac * n)
    if k > 0:
        lab, conf = model.predict(x)
        mark("infer")
        early[torch.topk(conf, k, largest=False).indices] = True
    side_rows = audit | early
    sidx = aidx if k == 0 else torch.nonzero(side_rows).squeeze(1)
    xs = x[sidx]
    # the audit's (and the earliest) full integrations overlap the network
    # passes and the short integrations
    side.wait_stream(main)
    with torch.cuda.stream(side):
        if stage_times is not None:
            # the audit launch's own span and step counts (for its roofline)
            a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a0.record(side)
            lab_side, audit_res = integrate_labels(solver, ode, xs, time_steady, with_result=True)
            a1.record(side)
        else:
            lab_side = integrate_labels(solver, ode, xs, time_steady)
    if k == 0:
        if stage_times is not None:
            i0 = torch.cuda.Event(enable_timing=True)
            i0.record(main)
        lab, conf = model.predict(x)
        mark("infer")
    routed = conf < gate
    evolved = torch.zeros_like(routed)
    lab_e = torch.full((n,), -1, dtype=torch.int32, device=dev)
    full = torch.full((n,), -1, dtype=torch.int32, device=dev)

    def evolve(rows: torch.Tensor) -> None:
        idx = torch.nonzero(rows).squeeze(1)
        if idx.numel() == 0:
            return
        lab2, conf2 = model.predict(evolve_states(solver, ode, x[idx], evolve_t))
        ok = conf2 >= gate
        # rows of one call are not evolved yet; index updates without a
        # device-to-host read
        evolved[idx] = ok
        lab_e[idx] = torch.where(ok, lab2, lab_e[idx])

    pending = []  # (row indices, labels) of full integrations issued on the main stream

    def integrate_async(rows: torch.Tensor) -> None:
        # on the main stream, so it overlaps the audit's stream rather than
        # queueing behind it
        idx = torch.nonzero(rows).squeeze(1)
        if idx.numel() == 0:
            return
        pending.append((idx, integrate_labels(solver, ode, x[idx], time_steady)))

    has_groups = group is not None and group_error_cap is not None

    def group_rate(pred, ref, valid):
        # disagreement of pred with ref per group, over audited rows where valid
        # (sums by index_add: bincount would read the largest index back)
        v = valid[aidx]
        wrong = ((pred[aidx] != ref[aidx]) & v).double()
        ga = gl[aidx]
        num = torch.zeros((ng,), dtype=torch.float64, device=dev).index_add_(0, ga, wrong)
        den = torch.zeros((ng,), dtype=torch.float64, device=dev).index_add_(0, ga, v.double())
        return num / den.clamp_min(1)

    done_full = early.clone()
    group_routed = torch.zeros_like(routed)
    distrusted = torch.zeros_like(routed)
    audit_evolved_error = None
    if evolve_t is not None:
        # short integrations of the below-gate rows and of the audit sample;
        # on the audit sample the first answer is checked against the answer
        # at the evolved state (self-consistency, known long before the
        # audit's full integrations end)
        evolve((routed | audit) & ~early)
        mark("evolve")
        if has_groups:
            bad_group = group_rate(lab, lab_e, audit & ~routed & evolved) > group_error_cap
            if bool(bad_group.any()):
                group_routed = bad_group[gl] & ~routed
                evolve(group_routed & ~audit & ~early)
        routed = routed | group_routed
        use_e = evolved & routed
        final = torch.where(use_e, lab_e, lab)
        first = routed & ~use_e & ~side_rows
        integrate_async(first)  # still below the gate: full span, now
        done_full |= first
        mark("late1")
    else:
        final = lab
        integrate_async(routed & ~side_rows)
        done_full |= routed & ~audit
        use_e = torch.zeros_like(routed)
    main.wait_stream(side)
    mark("audit")
    lab_side.record_stream(main)
    xs.record_stream(side)
    full[sidx] = lab_side
    if evolve_t is not None:
        ea = use_e & audit
        audit_evolved_error = float((lab_e[ea] != full[ea]).float().mean()) if bool(ea.any()) else None
    if has_groups:
        # the audit proper: a group whose labels so far disagree with its full
        # integrations on more than group_error_cap of the audited rows that
        # were not integrated is integrated entirely
        would_integrate = early | (routed & ~use_e)
        failed = group_rate(final, full, ~would_integrate) > group_error_cap
        distrusted = failed[gl]
        group_routed = group_routed | (distrusted & ~routed)
        routed = routed | distrusted
    mark("group")
    need_full = distrusted & ~audit & ~done_full
    integrate_async(need_full)
    mark("late2")
    n_late = sum(int(idx.numel()) for idx, _ in pending)
    for idx, lab_f in pending:
        full[idx] = lab_f
    # what the pass answers for the audited rows had they not been audited:
    # the integrator's label where it integrates them anyway, else its own
    integrates_anyway = need_full | done_full | distrusted | (routed & ~use_e)
    audit_final = torch.where(integrates_anyway, full, final)[aidx]
    integrated = need_full | done_full | audit
    final = torch.where(integrated, full, final)
    mark("end")
    if stage_times is not None:
        torch.cuda.synchronize(dev)
        for name, ev in marks[1:]:
            stage_times[name] = marks[0][1].elapsed_time(ev)
        stage_times["late_launches"] = len(pending)
        stage_times["audit_launch_ms"] = a0.elapsed_time(a1)
        stage_times["audit_nsteps"] = None if audit_res is None else audit_res.nsteps
        stage_times["audit_rows"] = int(xs.shape[0])
        if k == 0:
            stage_times["infer_launch_ms"] = i0.elapsed_time(marks[1][1])
    return GatedResult(labels=final, routed=routed, audit_idx=aidx, audit_surrogate=lab[aidx],
                       audit_integrated=full[aidx], integrated=integrated,
                       group_routed=group_routed, evolved=use_e & ~integrated,
                       audit_evolved_error=audit_evolved_error, n_late=n_late,
                       audit_final=audit_final)


def train_surrogate(solver, ode, x: torch.Tensor, n_train: int = 50_000, epochs: int = 600,
                    seed: int = 0, time_steady: float = 950.0, batch: int = 1024, lr: float = 3e-3):
    """Train a SurrogateMLP on n_train rows of x drawn at random (seeded) and
    labelled by the integrator (integrate_labels); returns (model, per-epoch
    loss, the training rows' indices)."""
    dev = solver.device
    n = x.shape[0]
    gsel = torch.Generator(device=dev).manual_seed(seed + 1)
    tidx = torch.randperm(n, generator=gsel, device=dev)[:n_train]
    ylab = integrate_labels(solver, ode, x[tidx], time_steady)
    model = SurrogateMLP(device=dev, seed=seed)
    hist = model.fit(x[tidx], ylab, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return model, hist, tidx


# ---- pendulum torque sweep (surrogate workload) ----------------------------

def torque_sweep_ics(T_values, counts, seed: int = 0, device="cuda:0"):
    """Rows (theta, omega, T) for every torque T with counts[i] ICs drawn
    uniformly in theta in [-pi + asin(T), pi + asin(T)], omega in [-10, 10]
    (the pendulum sampler region, case_studies/pendulum/setup_pendulum_system.py:28-32,
    with K = 1), plus the group index of every row."""
    g = torch.Generator().manual_seed(seed)
    rows, groups = [], []
    for gi, (T, c) in enumerate(zip(T_values, counts)):
        c = int(c)
        a = math.asin(min(1.0, float(T)))
        u = torch.rand(c, 2, generator=g)
        th = u[:, 0] * (2 * math.pi) + (-math.pi + a)
        om = u[:, 1] * 20.0 - 10.0
        rows.append(torch.stack([th, om, torch.full((c,), float(T))], 1))
        groups.append(torch.full((c,), gi, dtype=torch.int32))
    return torch.cat(rows).to(device), torch.cat(groups).to(device)


def basin_stability_by_group(labels: torch.Tensor, group: torch.Tensor, n_groups: int) -> torch.Tensor:
    """Fraction of label 0 (FP) per group."""
    cnt = torch.bincount(group.long(), minlength=n_groups).double()
    fp = torch.bincount(group.long(), weights=(labels == 0).double(), minlength=n_groups)
    return fp / cnt.clamp_min(1)


@dataclass
class SweepReport:
    n: int
    routed_frac: float
    audit_n: int
    audit_label_error: float
    max_bs_error_vs_plain: float | None
    bs_gated: list
    bs_plain: list | None
    train_loss: list
    train_n: int
    t_train_s: float
    t_gated_s: float
    t_plain_s: float | None


def run_torque_sweep(solver, ode, x, group, n_groups, *, train_n=50_000, gate=0.95,
                     audit_frac=0.02, epochs=60, seed=0, plain: bool = True,
                     time_steady: float = 950.0) -> SweepReport:
    """Train on train_n integrator-labelled rows, classify all rows through the
    gate, and (plain=True) integrate every row for the reference fractions."""
    import time

    dev = solver.device
    n = x.shape[0]
    gsel = torch.Generator(device=dev).manual_seed(seed + 1)
    tidx = torch.randperm(n, generator=gsel, device=dev)[:train_n]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ylab = integrate_labels(solver, ode, x[tidx], time_steady)
    model = SurrogateMLP(device=dev, seed=seed)
    hist = model.fit(x[tidx], ylab, epochs=epochs, batch=1024, lr=3e-3, seed=seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    t0 = time.perf_counter()
    res = classify_gated(model, solver, ode, x, gate=gate, audit_frac=audit_frac, group=group,
                         time_steady=time_steady, seed=seed)
    torch.cuda.synchronize()
    t_gated = time.perf_counter() - t0
    bs_g = basin_stability_by_group(res.labels, group, n_groups)
    aerr = float((res.audit_surrogate != res.audit_integrated).float().mean()) if len(res.audit_idx) else 0.0

    bs_p = None
    t_plain = None
    max_err = None
    if plain:
        t0 = time.perf_counter()
        plab = integrate_labels(solver, ode, x, time_steady)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        bs_p = basin_stability_by_group(plab, group, n_groups)
        max_err = float((bs_g - bs_p).abs().max())
    return SweepReport(n=n, routed_frac=float(res.routed.float().mean()), audit_n=len(res.audit_idx),
                       audit_label_error=aerr, max_bs_error_vs_plain=max_err,
                       bs_gated=bs_g.cpu().tolist(), bs_plain=None if bs_p is None else bs_p.cpu().tolist(),
                       train_loss=hist, train_n=train_n, t_train_s=t_train, t_gated_s=t_gated,
                       t_plain_s=t_plain)
