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


def sigma_derivative_table(mat, dt: float) -> np.ndarray:
    """Per material id: d ca, d cb / d sigma (float64 -> float32); ce does not
    depend on sigma.  With L = sigma dt/(2 eps): d ca/d sigma = -dt/(eps (1+L)^2),
    d cb/d sigma = -cb dt/(2 eps (1+L))."""
    eps_r, sig, _, _ = mat.props()
    eps = eps_r.astype(np.float64) * EPS_0
    L = sig.astype(np.float64) * dt / (2.0 * eps)
    cb = (dt / eps) / (1.0 + L)
    dca = -dt / (eps * (1.0 + L) ** 2)
    dcb = -cb * dt / (2.0 * eps * (1.0 + L))
    return np.ascontiguousarray(np.stack([dca, dcb], axis=1).astype(np.float32))


def dispersive_derivative_tables(prob, dt: float):
    """Derivative tables of a plan with Debye / Lorentz poles (float64 ->
    float32), replacing eps_derivative_table / sigma_derivative_table.

    With eps_inf = eps_r eps0, h = sigma dt/2, B = sum_d beta_d of the
    material's Debye poles, gamma = eps_inf + B + h and the dispersive ca =
    (eps_inf - B - h)/gamma, cb = dt/gamma (the coefficients of
    rfx_amd.materials.dispersion_tables, rfx/materials/debye.py:76-179 and
    lorentz.py:103-198; every pole-free material has B = 0 and the plain
    values): d ca/d eps_r = 2 eps0 (B + h)/gamma^2, d cb/d eps_r = -dt
    eps0/gamma^2, d ca/d sigma = -(dt/2)(gamma + eps_inf - B - h)/gamma^2,
    d cb/d sigma = -dt (dt/2)/gamma^2; ce = dt/(eps_r eps0) as without
    poles.  Every pole's cc is a constant over gamma, so the ADE part K of
    the E update has dK/dx = -(gamma'(x)/gamma) K: the third table holds
    (-eps0/gamma, -(dt/2)/gamma).  Returns (eps table [n_mat][3], sigma
    table [n_mat][2], K table [n_mat][2])."""
    from .materials import dispersion_tables
    mat = prob.mat
    _, _, rows = dispersion_tables(mat, dt, prob.poles, prob.lorentz)
    eps_r, sig, _, _ = mat.props()
    e = eps_r.astype(np.float64)
    einf = e * EPS_0
    h = sig.astype(np.float64) * dt / 2.0
    nd = len(prob.poles)
    B = rows[:, :nd, 1].astype(np.float64).sum(axis=1) if nd else np.zeros_like(e)
    g = einf + B + h
    num = einf - B - h
    ce = dt / einf
    d_eps = np.stack([2.0 * EPS_0 * (B + h) / g ** 2, -dt * EPS_0 / g ** 2, -ce / e], axis=1)
    d_sig = np.stack([-(dt / 2.0) * (g + num) / g ** 2, -dt * (dt / 2.0) / g ** 2], axis=1)
    d_k = np.stack([-EPS_0 / g, -(dt / 2.0) / g], axis=1)
    return tuple(np.ascontiguousarray(t.astype(np.float32)) for t in (d_eps, d_sig, d_k))


def mu_derivative_table(mat, dt: float) -> np.ndarray:
    """Per material id: d ch / d mu_r with ch = dt / (mu_r mu0), i.e. -ch/mu_r
    (float64 -> float32)."""
    _, _, mu_r, _ = mat.props()
    mu = mu_r.astype(np.float64)
    ch = dt / (mu * MU_0)
    return np.ascontiguousarray((-ch / mu).astype(np.float32))


class AdjointRunner:
    """Forward run with recording / checkpoints, then the reverse sweep."""

    def __init__(self, prob: Problem, n_steps: int, device, region=None,
                 checkpoint_segments: int | None = None,
                 record_budget: int = DEFAULT_RECORD_BUDGET, want_sigma: bool = False,
                 want_mu: bool = False, n_warmup: int = 0):
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
        self.disp = bool(prob.poles or prob.lorentz)
        per_step = (9 + (3 if want_mu else 0) + (3 if self.disp else 0)) * 4 * self.nreg
        state = 4 * (6 * int(np.prod(shape)) + int(self.dr.psis[0].numel()))
        if self.dr.pol is not None:
            state += 4 * int(self.dr.pol.numel())
        budget = int(record_budget)
        if self.dr.device.type == "cuda":
            free, _ = torch.cuda.mem_get_info(self.dr.device)
            budget = min(budget, int(0.8 * free))
        # steps [0, n_warmup) run without a record: the state after them is a
        # constant of the gradient (rfx/nonuniform.py:1734-1752)
        self.n_warmup = int(n_warmup)
        if not 0 <= self.n_warmup < max(self.n_steps, 1):
            raise ValueError(f"n_warmup ({self.n_warmup}) must be in [0, n_steps={self.n_steps})")
        n_opt = self.n_steps - self.n_warmup
        nseg = segment_count(n_opt, per_step, state, budget, checkpoint_segments)
        self.seg_len = math.ceil(n_opt / nseg)
        self.bounds = [(lo, min(self.n_steps, lo + self.seg_len))
                       for lo in range(self.n_warmup, self.n_steps, self.seg_len)]
        self.dcoef_disp = None
        if self.disp:
            self.dcoef, dsig, self.dcoef_disp = dispersive_derivative_tables(prob, prob.dt)
            self.dcoef_sigma = dsig if want_sigma else None
        else:
            self.dcoef = eps_derivative_table(prob.mat, prob.dt)
            self.dcoef_sigma = sigma_derivative_table(prob.mat, prob.dt) if want_sigma else None
        self.grad_sigma = None
        self.dcoef_mu = mu_derivative_table(prob.mat, prob.dt) if want_mu else None
        self.grad_mu = None
        self.rec_h = None
        self.rec_k = None
        self.rec = None
        self.checkpoints = []

    # -- forward -----------------------------------------------------------------
    def _record(self, lo, hi):
        rows = hi - lo
        if self.rec is None or self.rec.shape[0] < rows:
            self.rec = torch.empty((rows, 9, self.nreg), dtype=torch.float32,
                                   device=self.dr.device)
            if self.dcoef_mu is not None:
                self.rec_h = torch.empty((rows, 3, self.nreg), dtype=torch.float32,
                                         device=self.dr.device)
            if self.disp:
                self.rec_k = torch.empty((rows, 3, self.nreg), dtype=torch.float32,
                                         device=self.dr.device)
        rec_h = self.rec_h.data_ptr() if self.rec_h is not None else None
        rec_k = self.rec_k.data_ptr() if self.rec_k is not None else None
        with torch.cuda.device(self.dr.device):
            _lib.check(self.lib.rfxh_run_record_ex(
                self.dr.plan, lo, hi - lo, C.byref(self.region), self.rec.data_ptr(), rec_h,
                rec_k, self.dr.stream_ptr()))

    def forward(self) -> torch.Tensor:
        if self.n_warmup:
            self.dr.run(0, self.n_warmup, use_graph=True)
        if len(self.bounds) == 1:
            self._record(self.n_warmup, self.n_steps)
        else:
            for lo, hi in self.bounds:
                pol = self.dr.pol.clone() if self.dr.pol is not None else None
                self.checkpoints.append(([f.clone() for f in self.dr.fields], self.dr.psi.clone(),
                                         pol))
                self.dr.run(lo, hi - lo, use_graph=True)
        return self.dr.time_series().clone()

    def port_samples(self) -> torch.Tensor:
        """The lumped ports' (V, I) samples of the forward run, (T, n_port, 2)."""
        return self.dr.port_samples()[:self.n_steps].clone()

    def plane_values(self) -> list:
        """The DFT planes' accumulators of the forward run (complex128
        (n_freq, n1, n2) each)."""
        return [a.clone() for a in self.dr.plane_acc]

    def flux_values(self) -> list:
        """The flux monitors' accumulators, then the six faces of each NTFF
        box (complex128 (4, n_freq, n1, n2) each): the plan's flux order."""
        return [a.clone() for a in _flux_list(self.dr)]

    # -- reverse -------------------------------------------------------------------
    def backward(self, g_ts: torch.Tensor, g_ports: torch.Tensor | None = None,
                 g_planes: list | None = None, g_fluxes: list | None = None) -> torch.Tensor:
        """dJ/d eps_r on the whole grid from dJ/d(time series) and, with lumped
        ports, dJ/d(port samples) (T, n_port, 2), with DFT planes,
        dJ/d(accumulator) per plane, with flux monitors / NTFF boxes,
        dJ/d(accumulator) per entry of flux_values() (torch's complex
        gradient dJ/dRe + j dJ/dIm, or None); with want_sigma / want_mu
        also dJ/d sigma / dJ/d mu_r, left in ``grad_sigma`` / ``grad_mu``."""
        st = _ReverseState(self.prob, self.dr, self.n_steps, g_ts, self.region, self.nreg,
                           self.dcoef, self.dcoef_sigma, self.dcoef_mu, g_ports, g_planes,
                           g_fluxes)
        d = st.desc
        if self.dcoef_disp is not None:
            d.dcoef_disp = self.dcoef_disp.ctypes.data_as(C.POINTER(C.c_float))
        for s_idx in range(len(self.bounds) - 1, -1, -1):
            lo, hi = self.bounds[s_idx]
            if len(self.bounds) > 1:
                fields, psi, pol = self.checkpoints[s_idx]
                self.dr.load_state(fields, psi)
                if pol is not None:
                    self.dr.pol.copy_(pol)
                self._record(lo, hi)
            d.d_rec = self.rec.data_ptr()
            if self.rec_h is not None:
                d.d_rec_h = self.rec_h.data_ptr()
            if self.rec_k is not None:
                d.d_rec_k = self.rec_k.data_ptr()
            d.rec_step0 = lo
            d.rec_rows = hi - lo
            with torch.cuda.device(self.dr.device):
                _lib.check(self.lib.rfxh_adjoint(self.dr.plan, C.byref(d), lo, hi,
                                                 self.dr.stream_ptr()))
        torch.cuda.synchronize(self.dr.device)
        if self.dcoef_sigma is not None:
            self.grad_sigma = st.full_gradient(self.box, sigma=True)
        if self.dcoef_mu is not None:
            self.grad_mu = st.full_gradient(self.box, mu=True)
        return st.full_gradient(self.box)


class _ReverseState:
    """One instance's reverse-mode buffers (adjoint fields, adjoint CPML
    memory and its g-bar, dJ/dy of the probes, dJ/dw of the sources, the
    region gradient) and the descriptor that points at them."""

    def __init__(self, prob: Problem, dr: DeviceRun, T: int, g_ts: torch.Tensor, region,
                 nreg: int, dcoef: np.ndarray, dcoef_sigma: np.ndarray | None = None,
                 dcoef_mu: np.ndarray | None = None, g_ports: torch.Tensor | None = None,
                 g_planes: list | None = None, g_fluxes: list | None = None):
        dev = dr.device
        self.prob = prob
        self.T = T
        n_probe = len(prob.probes)
        self.probe_adj = torch.zeros((T, max(n_probe, 1)), dtype=torch.float32, device=dev)
        if n_probe and g_ts is not None:
            self.probe_adj[:, :n_probe] = g_ts.to(dev, torch.float32).reshape(T, n_probe)
        self.n_src = len(prob.sources)
        self.src_adj = torch.zeros((T, max(self.n_src, 1)), dtype=torch.float32, device=dev)
        npsi = int(dr.psis[0].numel())
        self.adj = [torch.zeros(prob.shape, dtype=torch.float32, device=dev) for _ in range(6)]
        self.adj_psi = torch.zeros(npsi, dtype=torch.float32, device=dev)
        self.gbar = torch.zeros(npsi, dtype=torch.float32, device=dev)
        self.grad = torch.zeros(nreg, dtype=torch.float32, device=dev)
        self.dcoef = dcoef
        d = _lib.AdjointDesc()
        for c in range(6):
            d.d_adj[c] = self.adj[c].data_ptr()
        d.d_adj_psi = self.adj_psi.data_ptr()
        d.d_gbar = self.gbar.data_ptr()
        d.d_probe_adj = self.probe_adj.data_ptr()
        d.d_src_adj = self.src_adj.data_ptr()
        d.region = region
        d.d_grad = self.grad.data_ptr()
        d.dcoef = self.dcoef.ctypes.data_as(C.POINTER(C.c_float))
        self.dcoef_sigma = dcoef_sigma
        self.grad_s = None
        if dcoef_sigma is not None:
            self.grad_s = torch.zeros(nreg, dtype=torch.float32, device=dev)
            d.d_grad_sigma = self.grad_s.data_ptr()
            d.dcoef_sigma = dcoef_sigma.ctypes.data_as(C.POINTER(C.c_float))
        self.dcoef_mu = dcoef_mu
        self.grad_m = None
        if dcoef_mu is not None:
            self.grad_m = torch.zeros(nreg, dtype=torch.float32, device=dev)
            d.d_grad_mu = self.grad_m.data_ptr()
            d.dcoef_mu = dcoef_mu.ctypes.data_as(C.POINTER(C.c_float))
        self.port_adj = None
        n_port = len(prob.ports)
        if n_port and g_ports is not None:
            # the plan's port table has max_steps rows: pad past T with zeros
            pa = torch.zeros((dr.port_out.shape[0], dr.port_out.shape[1], 2), dtype=torch.float32,
                             device=dev)
            pa[:T, :n_port] = g_ports.to(dev, torch.float32).reshape(T, n_port, 2)
            self.port_adj = pa
            d.d_port_adj = pa.data_ptr()
        self.pol_adj = None
        if dr.pol is not None:
            self.pol_adj = torch.zeros_like(dr.pol)
            d.d_pol_adj = self.pol_adj.data_ptr()
        self.acc_adj = []     # keeps the gradient tensors the descriptor points at
        self._ptr_tables = []
        if g_planes is not None and any(g is not None for g in g_planes):
            d.d_plane_adj = self._pointer_table(g_planes, dr.plane_acc, "plane", dev)
        if g_fluxes is not None and any(g is not None for g in g_fluxes):
            d.d_flux_adj = self._pointer_table(g_fluxes, _flux_list(dr), "flux", dev)
        self.desc = d

    def _pointer_table(self, grads, accs, what, dev):
        """Host array of device pointers to the complex128 gradients (NULL
        where a gradient is None), laid out like the accumulators."""
        if len(grads) != len(accs):
            raise ValueError(f"{len(grads)} {what} gradients for {len(accs)} accumulators")
        ptrs = (C.c_void_p * len(grads))()
        for q, (g, acc) in enumerate(zip(grads, accs)):
            if g is None:
                continue
            g = g.to(dev, torch.complex128).reshape(acc.shape).contiguous()
            self.acc_adj.append(g)
            ptrs[q] = g.data_ptr()
        self._ptr_tables.append(ptrs)
        return C.cast(ptrs, C.POINTER(C.c_void_p))

    def full_gradient(self, box, sigma: bool = False, mu: bool = False) -> torch.Tensor:
        """dJ/d eps_r (sigma: dJ/d sigma, mu: dJ/d mu_r) on the whole grid
        (zero outside the region box)."""
        prob = self.prob
        i0, i1, j0, j1, k0, k1 = box
        g = self.grad_m if mu else self.grad_s if sigma else self.grad
        full = torch.zeros(prob.shape, dtype=torch.float32, device=g.device)
        full[i0:i1, j0:j1, k0:k1] = g.reshape(i1 - i0, j1 - j0, k1 - k0)
        # uniform lane: source tables depend on eps_r through Cb at the
        # source cell (rfx/simulation.py:181-194, 275-284): w[n] = Cb * s[n].
        # The graded lane normalises sources with the materials before the
        # override (rfx/runners/nonuniform.py:630-643): no dependence.
        if self.n_src and not prob.graded and not mu:    # sources do not depend on mu_r
            sa = self.src_adj[:, :self.n_src].double().cpu().numpy()
            eps_r, sig, _, _ = prob.mat.props()
            for q, (idx, comp, wave) in enumerate(prob.sources):
                m = int(prob.mat.ids[tuple(idx)])
                e = float(eps_r[m])
                s = float(sig[m])
                eps = e * EPS_0
                L = s * prob.dt / (2.0 * eps)
                cb = (prob.dt / eps) / (1.0 + L)
                # d cb / d eps_r, or d cb / d sigma
                dcb = -cb * prob.dt / (2.0 * eps * (1.0 + L)) if sigma else -cb / (e * (1.0 + L))
                contrib = dcb / cb * float(np.dot(sa[:, q], np.asarray(wave[:self.T], np.float64)))
                full[tuple(idx)] += contrib
        return full


DEFAULT_BATCH_BUDGET = 192 << 30  # bytes: records and checkpoints of a batched reverse run


def batch_segment_steps(T, per_step, state, budget: int, forced: int | None = None) -> int:
    """Segment length S (steps) of a batched reverse run.

    Instance q (T[q] steps, record per_step[q] bytes per step, state
    checkpoint state[q] bytes) is cut into ceil(T[q]/S) segments; a
    segment's recomputation records all instances at once, so the
    device memory is  S * sum(per_step) + sum((ceil(T/S) - 1) * state)
    (the state at step 0 is the zero state, never stored).  S = max(T)
    (the forward records every step; no recomputation) when that fits,
    else the largest S that fits (fewest segments); MemoryError when no S
    does (fewer instances per batch, or smaller gradient regions, fit)."""
    T = np.asarray(T, np.int64)
    per = np.asarray(per_step, np.float64)
    st = np.asarray(state, np.float64)
    t_max = int(T.max())
    if forced is not None and int(forced) > 0:
        return max(1, min(int(forced), t_max))
    A = float(per.sum())
    if t_max * A <= budget:
        return t_max
    S = np.arange(1, t_max + 1, dtype=np.int64)
    need = S * A + ((np.ceil(T[None, :] / S[:, None]) - 1.0) * st[None, :]).sum(axis=1)
    ok = np.nonzero(need <= budget)[0]
    if len(ok) == 0:
        best = int(np.argmin(need))
        raise MemoryError(
            f"batched reverse mode needs at least {need[best] / 2**30:.1f} GiB (segments of "
            f"{best + 1} steps) but the budget is {budget / 2**30:.1f} GiB; split the batch "
            "or pass smaller gradient regions")
    return int(S[ok[-1]])


class BatchAdjointRunner:
    """Reverse mode of many instances together (include/rfxhip.h
    ``rfxh_batch_run_record`` / ``rfxh_batch_adjoint``), the batched
    counterpart of AdjointRunner: where the reference maps jax.grad over a
    sweep (rfx/vmap_sweep.py:1017-1022), every sub-step here is one launch
    over all instances still active.

    The instances' steps are cut into segments of S steps
    (``batch_segment_steps``).  The forward runs the batch segment by
    segment, keeping each instance's state at the segment starts; the
    backward, from the last segment to the first, restores every
    instance's state there, recomputes the segment for all of them with
    recording, and sweeps it in reverse for all of them.  With one segment
    the forward records directly and nothing is recomputed."""

    def __init__(self, problems, n_steps, device, regions=None,
                 record_budget: int = DEFAULT_BATCH_BUDGET, segment_steps: int | None = None):
        if not problems:
            raise ValueError("empty batch")
        self.problems = list(problems)
        n = len(self.problems)
        self.T = [int(v) for v in n_steps] if not np.isscalar(n_steps) else [int(n_steps)] * n
        if len(self.T) != n:
            raise ValueError("one step count per problem")
        for prob in self.problems:
            if prob.ports:
                raise NotImplementedError("the batched reverse mode covers probe objectives "
                                          "(no lumped-port S-parameters)")
        regions = [None] * n if regions is None else list(regions)
        self.runs = [DeviceRun(p, t, device, two_sets=False) for p, t in zip(self.problems, self.T)]
        self.device = self.runs[0].device
        self.lib = self.runs[0].lib
        self.boxes, self.regions, self.nreg = [], [], []
        for prob, reg in zip(self.problems, regions):
            shape = prob.shape
            if reg is None:
                reg = (0, shape[0], 0, shape[1], 0, shape[2])
            i0, i1, j0, j1, k0, k1 = (int(v) for v in reg)
            if not (0 <= i0 < i1 <= shape[0] and 0 <= j0 < j1 <= shape[1]
                    and 0 <= k0 < k1 <= shape[2]):
                raise ValueError(f"gradient region {reg} outside the grid {shape}")
            self.boxes.append((i0, i1, j0, j1, k0, k1))
            self.regions.append(_lib.Region(i0, j0, k0, i1 - i0, j1 - j0, k1 - k0))
            self.nreg.append((i1 - i0) * (j1 - j0) * (k1 - k0))
        per_step = [9 * 4 * v for v in self.nreg]
        state = [4 * (6 * int(np.prod(p.shape)) + int(r.psis[0].numel()))
                 for p, r in zip(self.problems, self.runs)]
        budget = int(record_budget)
        if self.device.type == "cuda":
            free, _ = torch.cuda.mem_get_info(self.device)
            budget = min(budget, int(0.8 * free))
        self.S = batch_segment_steps(self.T, per_step, state, budget, segment_steps)
        self.single = self.S >= max(self.T)
        self.bounds = [[(lo, min(t, lo + self.S)) for lo in range(0, t, self.S)] for t in self.T]
        self.dcoef = [eps_derivative_table(p.mat, p.dt) for p in self.problems]
        self.rec = [None] * n
        self.checkpoints = [dict() for _ in range(n)]
        plans = (C.c_void_p * n)(*[r.plan.value for r in self.runs])
        self.handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.rfxh_batch_create(plans, n, C.byref(self.handle)))

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _record(self, step0, counts, accumulate):
        n = len(self.runs)
        recs = (C.c_void_p * n)()
        for q in range(n):
            if counts[q] == 0:
                continue
            rows = min(self.S, self.T[q])
            if self.rec[q] is None:
                self.rec[q] = torch.empty((rows, 9, self.nreg[q]), dtype=torch.float32,
                                          device=self.device)
            recs[q] = self.rec[q].data_ptr()
        regs = (_lib.Region * n)(*self.regions)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.rfxh_batch_run_record(
                self.handle, (C.c_int64 * n)(*step0), (C.c_int64 * n)(*counts), regs, recs,
                1 if accumulate else 0, self._stream()))

    # -- forward -----------------------------------------------------------------
    def forward(self) -> list:
        n = len(self.runs)
        for r in self.runs:
            r.reset()
        if self.single:
            self._record([0] * n, self.T, True)
        else:
            for c in range(max(len(b) for b in self.bounds)):
                step0, counts = [], []
                for q, r in enumerate(self.runs):
                    if c < len(self.bounds[q]):
                        lo, hi = self.bounds[q][c]
                        if c > 0:
                            self.checkpoints[q][c] = ([f.clone() for f in r.fields], r.psi.clone())
                        step0.append(lo)
                        counts.append(hi - lo)
                    else:
                        step0.append(self.T[q])
                        counts.append(0)
                with torch.cuda.device(self.device):
                    _lib.check(self.lib.rfxh_batch_run(
                        self.handle, (C.c_int64 * n)(*step0), (C.c_int64 * n)(*counts),
                        self._stream(), 1))
        return [r.time_series().clone() for r in self.runs]

    # -- reverse -------------------------------------------------------------------
    def backward(self, g_ts) -> list:
        """g_ts[q]: dJ/d(time series of instance q); returns dJ/d eps_r per
        instance on its whole grid (zero outside its region)."""
        n = len(self.runs)
        states = [_ReverseState(p, r, t, g, reg, nr, dc) for p, r, t, g, reg, nr, dc in
                  zip(self.problems, self.runs, self.T, g_ts, self.regions, self.nreg, self.dcoef)]
        for c in range(max(len(b) for b in self.bounds) - 1, -1, -1):
            part = [q for q in range(n) if c < len(self.bounds[q])]
            lo = [0] * n
            hi = [0] * n
            for q in part:
                lo[q], hi[q] = self.bounds[q][c]
            if not self.single:
                for q in part:
                    r = self.runs[q]
                    if c == 0:
                        r.reset()
                    else:
                        r.load_state(*self.checkpoints[q][c])
                self._record(lo, [h - l for l, h in zip(lo, hi)], False)
            descs = (C.POINTER(_lib.AdjointDesc) * n)()
            for q in part:
                d = states[q].desc
                d.d_rec = self.rec[q].data_ptr()
                d.rec_step0 = lo[q]
                d.rec_rows = hi[q] - lo[q]
                descs[q] = C.pointer(d)
            with torch.cuda.device(self.device):
                _lib.check(self.lib.rfxh_batch_adjoint(
                    self.handle, descs, (C.c_int64 * n)(*lo), (C.c_int64 * n)(*hi),
                    self._stream()))
        torch.cuda.synchronize(self.device)
        return [st.full_gradient(box) for st, box in zip(states, self.boxes)]

    def close(self):
        if self.handle:
            self.lib.rfxh_batch_destroy(self.handle)
            self.handle = C.c_void_p()
        for r in self.runs:
            r.close()
        self.rec = [None] * len(self.runs)
        self.checkpoints = [dict() for _ in self.runs]

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_gradients(problems, n_steps, dj_dts, device, regions=None,
                    record_budget: int = DEFAULT_BATCH_BUDGET):
    """Time series and dJ/d eps_r of many instances by batched reverse mode.

    dj_dts(q, ts) returns dJ/d(time series) of instance q from its time
    series.  The instances run as one BatchAdjointRunner when its records
    and checkpoints fit the budget, else as consecutive halves (recursively).
    Returns (time series, gradients), one of each per instance, in order."""
    probs = list(problems)
    steps = [int(v) for v in n_steps] if not np.isscalar(n_steps) else [int(n_steps)] * len(probs)
    regs = [None] * len(probs) if regions is None else list(regions)

    def run(lo, hi):
        try:
            r = BatchAdjointRunner(probs[lo:hi], steps[lo:hi], device, regs[lo:hi], record_budget)
        except MemoryError:
            if hi - lo == 1:
                raise
            mid = (lo + hi) // 2
            a, b = run(lo, mid), run(mid, hi)
            return a[0] + b[0], a[1] + b[1]
        try:
            ts = r.forward()
            grads = r.backward([dj_dts(lo + q, t) for q, t in enumerate(ts)])
            return ts, grads
        finally:
            r.close()

    return run(0, len(probs))


def _flux_list(dr) -> list:
    """A run's flux accumulators in the plan's order: the flux monitors,
    then the six faces of each NTFF box."""
    return list(dr.flux_acc) + [f for box in dr.ntff_acc for f in box]


class _ForwardFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, mu, runner):
        ctx.runner = runner
        ctx.shapes = tuple(None if t is None else (t.shape, t.device) for t in (eps, sigma, mu))
        ctx.set_materialize_grads(False)   # outputs the objective ignores cost nothing
        ts = runner.forward()
        planes = runner.plane_values()
        ctx.n_plane = len(planes)
        return (ts, runner.port_samples(), *planes, *runner.flux_values())

    @staticmethod
    def backward(ctx, g_ts, g_ports, *g_acc):
        r = ctx.runner
        k = ctx.n_plane
        grads = (r.backward(g_ts, g_ports if r.prob.ports else None, list(g_acc[:k]),
                            list(g_acc[k:])),
                 r.grad_sigma, r.grad_mu)
        out = [None if sh is None or w is None else w.reshape(sh[0]).to(sh[1])
               for w, sh in zip(grads, ctx.shapes)]
        return out[0], out[1], out[2], None


def forward_with_grad(sim, prob: Problem, grid, eps_tensor, n_steps: int,
                      checkpoint_segments=None, grad_region=None, sigma_tensor=None,
                      mu_tensor=None, n_warmup: int = 0):
    """The time series, the lumped ports' (V, I) samples (T, n_port, 2), the
    DFT planes' accumulators and the flux accumulators (flux monitors, then
    NTFF faces) as functions of eps_override, sigma_override and
    mu_r_override (the tensors among them that require grad)."""
    def req(t):
        return t if isinstance(t, torch.Tensor) and t.requires_grad else None

    eps_in, sig_in, mu_in = req(eps_tensor), req(sigma_tensor), req(mu_tensor)
    runner = AdjointRunner(prob, n_steps, sim._device, grad_region, checkpoint_segments,
                           want_sigma=sig_in is not None, want_mu=mu_in is not None,
                           n_warmup=n_warmup)
    ts, ports, *acc = _ForwardFn.apply(eps_in, sig_in, mu_in, runner)
    k = len(prob.planes)
    return ts, ports, acc[:k], acc[k:], runner.dr
