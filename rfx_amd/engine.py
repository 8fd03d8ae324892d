"""Device state and plan for one FDTD problem.

``Problem`` is the host-side description produced by the set-up code
(grid, material ids, CPML tables, source tables, probes, ports, DFT
planes).  ``DeviceRun`` uploads it into torch tensors on the GPU, creates
an rfxhip plan over those buffers and advances it.  Everything that scales
with the grid lives in HBM; the host keeps only the small tables.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .materials import MaterialMap

COMP_INDEX = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}


@dataclass
class Problem:
    shape: tuple
    dt: float
    graded: bool
    inv_e: list                      # 3 float32 arrays
    inv_h: list
    mat: MaterialMap
    faces: list                      # 6 x (n, b, c, km1, inv_d)
    pec_axes: str = "xyz"
    sources: list = field(default_factory=list)   # (idx, comp, wave f32[n_steps])
    probes: list = field(default_factory=list)    # (idx, comp)
    ports: list = field(default_factory=list)     # (idx, comp, d_par, (lx, ly, lz))
    planes: list = field(default_factory=list)    # (axis, index, comp, freqs f64)
    # flux monitors: (axis, index, lo1, hi1, lo2, hi2, freqs f64, weights f32 [n+1] | None)
    fluxes: list = field(default_factory=list)
    record_order: str = "uniform"                 # "uniform" | "graded"
    # near-to-far-field boxes: (lo (i, j, k), hi (i, j, k), freqs f64)
    ntff: list = field(default_factory=list)
    # dispersion poles: Debye (delta_eps, tau) as poles 0.., then Lorentz
    # (omega_0, delta, kappa); pole q acts where the material flags carry bit 16 + q
    poles: list = field(default_factory=list)
    lorentz: list = field(default_factory=list)

    @property
    def plane_t_offset(self):
        # uniform: t = (n+1) dt (rfx/simulation.py:1575); graded: t = n dt
        return 1 if self.record_order == "uniform" else 0


def _cell(shape, idx):
    return (int(idx[0]) * shape[1] + int(idx[1])) * shape[2] + int(idx[2])


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class DeviceRun:
    """Device buffers + rfxhip plan for ``n_steps`` steps of one problem."""

    def __init__(self, prob: Problem, n_steps: int, device=None, two_sets: bool = True):
        _lib.require_device()
        self.lib = _lib.load()
        self.prob = prob
        self.n_steps = int(n_steps)
        self.device = torch.device(device if device is not None else "cuda")
        dev = self.device
        shape = prob.shape
        nx, ny, nz = shape
        ncell = nx * ny * nz
        rank = {"uniform": {_lib.OP_PORT: 0, _lib.OP_INJECT: 1, _lib.OP_PROBE: 2},
                "graded": {_lib.OP_INJECT: 0, _lib.OP_PORT: 1, _lib.OP_PROBE: 2}}[prob.record_order]
        ops = []
        for s, (idx, comp, _) in enumerate(prob.sources):
            ops.append((_cell(shape, idx), _lib.OP_INJECT, COMP_INDEX[comp], s, 0.0, (0, 0, 0), idx))
        for s, (idx, comp) in enumerate(prob.probes):
            ops.append((_cell(shape, idx), _lib.OP_PROBE, COMP_INDEX[comp], s, 0.0, (0, 0, 0), idx))
        for s, (idx, comp, dpar, loop) in enumerate(prob.ports):
            ops.append((_cell(shape, idx), _lib.OP_PORT, COMP_INDEX[comp], s, dpar, loop, idx))
        # cells with point operations get the POINT flag (a variant material id)
        mat = prob.mat
        for op in ops:
            mat.set_cell(tuple(op[6]), add_flags=_lib.RFXH_POINT)
        if prob.poles or prob.lorentz:
            from .materials import POLE_SHIFT, dispersion_tables
            self.coef, self.flags, self.pole_rows = dispersion_tables(mat, prob.dt, prob.poles,
                                                                      prob.lorentz)
            # the box holding every cell with poles carries the polarisation
            disp = np.nonzero((self.flags >> POLE_SHIFT) != 0)[0]
            cells = np.isin(mat.ids, disp.astype(np.uint16))
            lo, n = [], []
            for a in range(3):
                any_a = np.any(cells, axis=tuple(b for b in range(3) if b != a))
                idx = np.nonzero(any_a)[0]
                lo.append(int(idx[0]))
                n.append(int(idx[-1]) - int(idx[0]) + 1)
            self.pol_box = (tuple(lo), tuple(n))
        else:
            self.coef, self.flags = mat.coefficient_table(prob.dt)
            self.pole_rows, self.pol_box = None, None
        # subpixel smoothing: Ey, Ez coefficients of the per-component ids
        self.acoef = mat.aniso_table(prob.dt) if not (prob.poles or prob.lorentz) else None
        self._ops = (_lib.PointOp * max(len(ops), 1))()
        for q, (cell, kind, comp, slot, dpar, loop, _) in enumerate(ops):
            o = self._ops[q]
            o.cell, o.kind, o.comp, o.slot = cell, kind, comp, slot
            o.order = rank[kind] * 1_000_000 + slot
            o.d_par = dpar
            o.loop[0], o.loop[1], o.loop[2] = loop
        n_ops = len(ops)

        # ---- device buffers -------------------------------------------------
        self.ids = torch.from_numpy(mat.ids.view(np.int16)).to(dev)
        # two field sets: each step reads one and writes the other (one fused
        # kernel per step); one set: in-place two-pass step
        self.sets = [[torch.zeros(shape, dtype=torch.float32, device=dev) for _ in range(6)]]
        if two_sets:
            self.sets.append([torch.zeros(shape, dtype=torch.float32, device=dev)
                              for _ in range(6)])
        n_src, n_probe, n_port = len(prob.sources), len(prob.probes), len(prob.ports)
        T = self.n_steps
        if n_src:
            w = np.stack([np.asarray(s[2][:T], np.float32) for s in prob.sources], axis=1)
            if w.shape[0] < T:
                raise ValueError("source table shorter than n_steps")
            self.src_wave = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
        else:
            self.src_wave = None
        self.probe_out = torch.zeros((T, max(n_probe, 1)), dtype=torch.float32, device=dev)
        self.pol = None
        if prob.poles or prob.lorentz:
            nbox = int(np.prod(self.pol_box[1]))
            n_state = 3 * len(prob.poles) + 6 * len(prob.lorentz)
            self.pol = torch.zeros(n_state * nbox, dtype=torch.float32, device=dev)
        self.port_out = torch.zeros((T, max(n_port, 1), 2), dtype=torch.float32, device=dev)
        self.plane_acc = []
        planes_c = (_lib.Plane * max(len(prob.planes), 1))()
        self._plane_freqs = []
        for q, (axis, index, comp, freqs) in enumerate(prob.planes):
            n1 = shape[1] if axis == 0 else shape[0]
            n2 = shape[1] if axis == 2 else shape[2]
            fr = np.ascontiguousarray(np.asarray(freqs, np.float64))
            self._plane_freqs.append(fr)
            acc = torch.zeros((len(fr), n1, n2), dtype=torch.complex128, device=dev)
            self.plane_acc.append(acc)
            pc = planes_c[q]
            pc.axis, pc.index, pc.comp, pc.n_freq = axis, int(index), COMP_INDEX[comp], len(fr)
            pc.freqs = fr.ctypes.data_as(C.POINTER(C.c_double))
            pc.d_acc = acc.data_ptr()
        self._planes_c = planes_c
        self.flux_acc = []
        flux_c = (_lib.Flux * max(len(prob.fluxes), 1))()
        for q, (axis, index, lo1, hi1, lo2, hi2, freqs, wts) in enumerate(prob.fluxes):
            fr = np.ascontiguousarray(np.asarray(freqs, np.float64))
            self._plane_freqs.append(fr)
            acc = torch.zeros((4, len(fr), hi1 - lo1, hi2 - lo2), dtype=torch.complex128, device=dev)
            self.flux_acc.append(acc)
            fc = flux_c[q]
            fc.axis, fc.index = axis, int(index)
            fc.lo1, fc.n1, fc.lo2, fc.n2 = lo1, hi1 - lo1, lo2, hi2 - lo2
            fc.n_freq = len(fr)
            fc.freqs = fr.ctypes.data_as(C.POINTER(C.c_double))
            if wts is not None:
                w = np.ascontiguousarray(np.asarray(wts, np.float32))
                if len(w) < T + 1:
                    raise ValueError("flux window table shorter than n_steps + 1")
                self._plane_freqs.append(w)
                fc.weight = _f32p(w)
            fc.d_acc = acc.data_ptr()
        self._flux_c = flux_c
        # NTFF boxes: six face accumulators [4][n_freqs][n1][n2] each
        self.ntff_acc = []
        ntff_c = (_lib.Ntff * max(len(prob.ntff), 1))()
        for q, (lo, hi, freqs) in enumerate(prob.ntff):
            fr = np.ascontiguousarray(np.asarray(freqs, np.float64))
            self._plane_freqs.append(fr)
            ni, nj, nk = (hi[a] - lo[a] for a in range(3))
            faces = []
            for n1, n2 in ((nj, nk), (nj, nk), (ni, nk), (ni, nk), (ni, nj), (ni, nj)):
                faces.append(torch.zeros((4, len(fr), n1, n2), dtype=torch.complex128, device=dev))
            self.ntff_acc.append(faces)
            nc = ntff_c[q]
            for a in range(3):
                nc.lo[a], nc.hi[a] = int(lo[a]), int(hi[a])
            nc.n_freq = len(fr)
            nc.freqs = fr.ctypes.data_as(C.POINTER(C.c_double))
            for f in range(6):
                nc.d_acc[f] = faces[f].data_ptr()
        self._ntff_c = ntff_c

        # ---- descriptor -----------------------------------------------------
        d = _lib.Desc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.graded = 1 if prob.graded else 0
        self._keep = []
        for a in range(3):
            e = np.ascontiguousarray(prob.inv_e[a], np.float32)
            h = np.ascontiguousarray(prob.inv_h[a], np.float32)
            self._keep += [e, h]
            d.inv_e[a] = _f32p(e)
            d.inv_h[a] = _f32p(h)
        d.dt = prob.dt
        d.d_mat = self.ids.data_ptr()
        d.n_mat = len(self.flags)
        d.mat_coef = _f32p(self.coef)
        d.mat_flags = self.flags.ctypes.data_as(C.POINTER(C.c_uint32))
        for f, (n, b, c, km1, inv_d) in enumerate(prob.faces):
            d.cpml_n[f] = n
            if n:
                arrs = [np.ascontiguousarray(x, np.float32) for x in (b, c, km1)]
                self._keep += arrs
                d.cpml_b[f], d.cpml_c[f], d.cpml_km1[f] = (_f32p(x) for x in arrs)
                d.cpml_inv_d[f] = inv_d
        d.pec_axes = sum(1 << a for a in range(3) if "xyz"[a] in prob.pec_axes)
        for c in range(6):
            d.d_f[c] = self.sets[0][c].data_ptr()
            if two_sets:
                d.d_f2[c] = self.sets[1][c].data_ptr()
        npsi = self.lib.rfxh_psi_floats(C.byref(d))
        self.psis = [torch.zeros(max(int(npsi), 1), dtype=torch.float32, device=dev)
                     for _ in range(len(self.sets))]
        d.d_psi = self.psis[0].data_ptr()
        if two_sets:
            d.d_psi2 = self.psis[1].data_ptr()
        d.n_ops = n_ops
        d.ops = self._ops
        d.n_src = n_src
        d.d_src_wave = self.src_wave.data_ptr() if self.src_wave is not None else None
        d.n_probe = n_probe
        d.d_probe_out = self.probe_out.data_ptr()
        d.n_port = n_port
        d.d_port_out = self.port_out.data_ptr()
        d.max_steps = T
        d.n_plane = len(prob.planes)
        d.planes = planes_c
        d.plane_t_offset = prob.plane_t_offset
        d.n_flux = len(prob.fluxes)
        d.flux = flux_c
        d.n_ntff = len(prob.ntff)
        d.ntff = ntff_c
        if prob.poles or prob.lorentz:
            d.n_pole = len(prob.poles) + len(prob.lorentz)
            d.n_debye = len(prob.poles)
            d.pole_coef = _f32p(self.pole_rows)
            for a in range(3):
                d.pol_lo[a] = self.pol_box[0][a]
                d.pol_n[a] = self.pol_box[1][a]
            d.d_pol = self.pol.data_ptr()
        if self.acoef is not None:
            d.aniso_coef = _f32p(self.acoef)
        self.desc = d
        self.plan = C.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(self.lib.rfxh_plan_create(C.byref(d), C.byref(self.plan)))

    # ------------------------------------------------------------------------
    def stream_ptr(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def run(self, step0: int = 0, n_steps: int | None = None, use_graph: bool = True):
        n = self.n_steps - step0 if n_steps is None else int(n_steps)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.rfxh_run(self.plan, int(step0), n, self.stream_ptr(),
                                         1 if use_graph else 0))

    def state_set(self) -> int:
        return int(self.lib.rfxh_state_set(self.plan))

    @property
    def fields(self):
        """The six field tensors (ex, ey, ez, hx, hy, hz) holding the current state."""
        return self.sets[self.state_set()]

    @property
    def psi(self):
        return self.psis[self.state_set()]

    def load_state(self, fields, psi):
        """Write a state (six field tensors and the CPML memory) into set 0
        and make it current."""
        for dst, src in zip(self.sets[0], fields):
            dst.copy_(src)
        self.psis[0].copy_(psi)
        _lib.check(self.lib.rfxh_set_state_set(self.plan, 0))

    def reset(self):
        """Zero the state and outputs (a fresh run over the same problem)."""
        for f in self.sets[0]:
            f.zero_()
        self.psis[0].zero_()
        _lib.check(self.lib.rfxh_set_state_set(self.plan, 0))
        self.probe_out.zero_()
        self.port_out.zero_()
        for a in self.plane_acc:
            a.zero_()
        for a in self.flux_acc:
            a.zero_()
        for faces in self.ntff_acc:
            for a in faces:
                a.zero_()
        if self.pol is not None:
            self.pol.zero_()

    def set_timing(self, enable: bool):
        _lib.check(self.lib.rfxh_set_timing(self.plan, 1 if enable else 0))

    def kernel_time_ms(self, which: int):
        avg = C.c_double()
        cnt = C.c_int64()
        _lib.check(self.lib.rfxh_kernel_time_ms(self.plan, which, C.byref(avg), C.byref(cnt)))
        return avg.value, cnt.value

    def close(self):
        if self.plan:
            self.lib.rfxh_plan_destroy(self.plan)
            self.plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- outputs --------------------------------------------------------------
    def time_series(self):
        return self.probe_out[:, :len(self.prob.probes)]

    def port_samples(self):
        return self.port_out[:, :len(self.prob.ports), :]

    def field(self, comp):
        return self.fields[COMP_INDEX[comp]]
