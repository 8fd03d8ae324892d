"""Reverse-mode gradient of the FDTD forward run w.r.t. eps_r.

Stands in for ``jax.grad`` through the reference's ``lax.scan``
(rfx/api/_execute.py:2447-2858; per-step ``jax.checkpoint`` at
rfx/simulation.py:1975 and sqrt(n) segments at :2011-2055).  The HIP side
provides the transposed step (include/rfxhip.h ``rfxh_adjoint``) and a
forward record of what dE/d(ca, cb, ce) needs (``rfxh_run_record``).

Memory: the record takes 9 floats per region cell per step.  When all steps
fit the budget the forward records them in one go and the reverse sweep
runs once (forward + adjoint, no recomputation).  Otherwise the run is cut
into segments: the forward keeps one state checkpoint (fields + CPML
memory) per segment, and the backward restores each checkpoint, recomputes
its segment with recording, and sweeps it in reverse (latest segment
first).  ``checkpoint_segments=K`` forces K segments.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .constants import EPS_0, MU_0
from .engine import DeviceRun, Problem

DEFAULT_RECORD_BUDGET = 24 << 30  # bytes of device memory for the record and checkpoints


def segment_count(n_steps: int, per_step: int, state: int, budget: int,
                  forced: int | None = None) -> int:
    """Number of recompute segments of a reverse-mode run.

    One segment (no checkpoints, no recompute) when the record of every
    step fits the budget.  Otherwise K segments cost a record of
    ceil(T/K) steps plus K state checkpoints (fields + CPML memory); the
    total is smallest near K = sqrt(T per_step / state), the reference's
    sqrt(n) split (rfx/simulation.py:2011-2055).  The run's cost does not
    depend on K beyond K = 1 (every step is recomputed once), so the
    smallest K whose memory fits is taken.  Raises MemoryError when even
    the best K does not fit (a smaller grad_region shrinks the record)."""
    T = int(n_steps)
    if forced is not None and int(forced) > 1:
        return min(int(forced), max(T, 1))
    if T * per_step <= budget:
        return 1

    def need(k):
        return math.ceil(T / k) * per_step + k * state

    k_best = max(2, min(T, int(round(math.sqrt(T * per_step / max(state, 1))))))
    k_best = min((k for k in (k_best - 1, k_best, k_best + 1) if 2 <= k <= T), key=need)
    if need(k_best) > budget:
        raise MemoryError(
            f"reverse mode needs at least {need(k_best) / 2**30:.1f} GiB ({k_best} segments of "
            f"{math.ceil(T / k_best)} steps: record {per_step} B/step, checkpoint {state} B) "
            f"but the budget is {budget / 2**30:.1f} GiB; pass a smaller grad_region")
    lo, hi = 2, k_best                    # need() falls from 2 to k_best: bisect the first fit
    while lo < hi:
        mid = (lo + hi) // 2
        if need(mid) <= budget:
            hi = mid
        else:
            lo = mid + 1
    return lo


def eps_derivative_table(mat, dt: float) -> np.ndarray:
    """Per material id: d ca, d cb, d ce / d eps_r (float64 -> float32).

    ca = (1-L)/(1+L), cb = (dt/eps)/(1+L), L = sigma dt/(2 eps), eps = eps_r eps0
    (rfx/core/yee.py:386-393); ce = dt/(eps_r eps0) (rfx/boundaries/cpml.py:538).
    """
    eps_r, sig, _, _ = mat.props()
    e = eps_r.astype(np.float64)
    s = sig.astype(np.float64)
    eps = e * EPS_0
    L = s * dt / (2.0 * eps)
    cb = (dt / eps) / (1.0 + L)
    ce = dt / eps
    dca = 2.0 * L / (e * (1.0 + L) ** 2)
    dcb = -cb / (e * (1.0 + L))
    dce = -ce / e
    return np.ascontiguousarray(np.stack([dca, dcb, dce], axis=1).astype(np.float32))


class AdjointRunner:
    """Forward run with recording / checkpoints, then the reverse sweep."""

    def __init__(self, prob: Problem, n_steps: int, device, region=None,
                 checkpoint_segments: int | None = None,
                 record_budget: int = DEFAULT_RECORD_BUDGET):
        if prob.ports:
            raise NotImplementedError("gradients through lumped-port V/I samples are not supported")
        self.prob = prob
        self.n_steps = int(n_steps)
        self.dr = DeviceRun(prob, n_steps, device, two_sets=False)
        self.lib = self.dr.lib
        shape = prob.shape
        if region is None:
            region = (0, shape[0], 0, shape[1], 0, shape[2])
        i0, i1, j0, j1, k0, k1 = (int(v) for v in region)
        if not (0 <= i0 < i1 <= shape[0] and 0 <= j0 < j1 <= shape[1] and 0 <= k0 < k1 <= shape[2]):
            raise ValueError(f"gradient region {region} outside the grid {shape}")
        self.box = (i0, i1, j0, j1, k0, k1)
        self.region = _lib.Region(i0, j0, k0, i1 - i0, j1 - j0, k1 - k0)
        self.nreg = (i1 - i0) * (j1 - j0) * (k1 - k0)
        per_step = 9 * 4 * self.nreg
        state = 4 * (6 * int(np.prod(shape)) + int(self.dr.psis[0].numel()))
        budget = int(record_budget)
        if self.dr.device.type == "cuda":
            free, _ = torch.cuda.mem_get_info(self.dr.device)
            budget = min(budget, int(0.8 * free))
        nseg = segment_count(self.n_steps, per_step, state, budget, checkpoint_segments)
        self.seg_len = math.ceil(self.n_steps / nseg)
        self.bounds = [(lo, min(self.n_steps, lo + self.seg_len))
                       for lo in range(0, self.n_steps, self.seg_len)]
        self.dcoef = eps_derivative_table(prob.mat, prob.dt)
        self.rec = None
        self.checkpoints = []

    # -- forward -----------------------------------------------------------------
    def _record(self, lo, hi):
        rows = hi - lo
        if self.rec is None or self.rec.shape[0] < rows:
            self.rec = torch.empty((rows, 9, self.nreg), dtype=torch.float32,
                                   device=self.dr.device)
        with torch.cuda.device(self.dr.device):
            _lib.check(self.lib.rfxh_run_record(
                self.dr.plan, lo, hi - lo, C.byref(self.region), self.rec.data_ptr(),
                self.dr.stream_ptr()))

    def forward(self) -> torch.Tensor:
        if len(self.bounds) == 1:
            self._record(0, self.n_steps)
        else:
            for lo, hi in self.bounds:
                self.checkpoints.append(([f.clone() for f in self.dr.fields], self.dr.psi.clone()))
                self.dr.run(lo, hi - lo, use_graph=True)
        return self.dr.time_series().clone()

    # -- reverse -------------------------------------------------------------------
    def backward(self, g_ts: torch.Tensor) -> torch.Tensor:
        dev = self.dr.device
        n_probe = len(self.prob.probes)
        T = self.n_steps
        probe_adj = torch.zeros((T, max(n_probe, 1)), dtype=torch.float32, device=dev)
        if n_probe:
            probe_adj[:, :n_probe] = g_ts.to(dev, torch.float32).reshape(T, n_probe)
        n_src = len(self.prob.sources)
        src_adj = torch.zeros((T, max(n_src, 1)), dtype=torch.float32, device=dev)
        npsi = int(self.dr.psis[0].numel())
        adj = [torch.zeros(self.prob.shape, dtype=torch.float32, device=dev) for _ in range(6)]
        adj_psi = torch.zeros(npsi, dtype=torch.float32, device=dev)
        gbar = torch.zeros(npsi, dtype=torch.float32, device=dev)
        grad = torch.zeros(self.nreg, dtype=torch.float32, device=dev)
        d = _lib.AdjointDesc()
        for c in range(6):
            d.d_adj[c] = adj[c].data_ptr()
        d.d_adj_psi = adj_psi.data_ptr()
        d.d_gbar = gbar.data_ptr()
        d.d_probe_adj = probe_adj.data_ptr()
        d.d_src_adj = src_adj.data_ptr()
        d.region = self.region
        d.d_grad = grad.data_ptr()
        d.dcoef = self.dcoef.ctypes.data_as(C.POINTER(C.c_float))
        for s_idx in range(len(self.bounds) - 1, -1, -1):
            lo, hi = self.bounds[s_idx]
            if len(self.bounds) > 1:
                fields, psi = self.checkpoints[s_idx]
                self.dr.load_state(fields, psi)
                self._record(lo, hi)
            d.d_rec = self.rec.data_ptr()
            d.rec_step0 = lo
            d.rec_rows = hi - lo
            with torch.cuda.device(dev):
                _lib.check(self.lib.rfxh_adjoint(self.dr.plan, C.byref(d), lo, hi,
                                                 self.dr.stream_ptr()))
        torch.cuda.synchronize(dev)
        i0, i1, j0, j1, k0, k1 = self.box
        full = torch.zeros(self.prob.shape, dtype=torch.float32, device=dev)
        full[i0:i1, j0:j1, k0:k1] = grad.reshape(i1 - i0, j1 - j0, k1 - k0)
        # uniform lane: source tables depend on eps_r through Cb at the
        # source cell (rfx/simulation.py:181-194, 275-284): w[n] = Cb * s[n].
        # The graded lane normalises sources with the materials before the
        # override (rfx/runners/nonuniform.py:630-643): no dependence.
        if n_src and not self.prob.graded:
            sa = src_adj[:, :n_src].double().cpu().numpy()
            eps_r, sig, _, _ = self.prob.mat.props()
            for q, (idx, comp, wave) in enumerate(self.prob.sources):
                m = int(self.prob.mat.ids[tuple(idx)])
                e = float(eps_r[m])
                s = float(sig[m])
                eps = e * EPS_0
                L = s * self.prob.dt / (2.0 * eps)
                cb = (self.prob.dt / eps) / (1.0 + L)
                dcb = -cb / (e * (1.0 + L))
                contrib = dcb / cb * float(np.dot(sa[:, q], np.asarray(wave[:T], np.float64)))
                full[tuple(idx)] += contrib
        return full


class _ForwardFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, runner):
        ctx.runner = runner
        ctx.eps_device = eps.device
        ctx.eps_shape = eps.shape
        return runner.forward()

    @staticmethod
    def backward(ctx, g):
        grad = ctx.runner.backward(g)
        return grad.reshape(ctx.eps_shape).to(ctx.eps_device), None


def forward_with_grad(sim, prob: Problem, grid, eps_tensor: torch.Tensor, n_steps: int,
                      checkpoint_segments=None, grad_region=None):
    runner = AdjointRunner(prob, n_steps, sim._device, grad_region, checkpoint_segments)
    ts = _ForwardFn.apply(eps_tensor, runner)
    return ts, runner.dr
