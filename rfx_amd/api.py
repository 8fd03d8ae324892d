"""``Simulation`` facade: the reference's user API over the HIP backend.

Mirrors rfx/api/__init__.py:280-2600 (builder), rfx/api/_compile.py:60-367
(grid and materials), rfx/runners/uniform.py:140-760 (run lane) and
rfx/api/_execute.py:917-1669, 2447-2858 (forward lane) for the FDTD
time-stepping path: uniform and graded Yee grids, CPML, PEC faces and PEC
geometry masks, soft sources, lumped ports, point probes, DFT planes and
lumped-port S11.  Everything else the reference offers raises
NotImplementedError instead of running a different algorithm.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np
import torch

from .constants import EPS_0
from .boundary import BoundarySpec, run_kind
from .cpml import face_tables
from .engine import DeviceRun, Problem
from .geometry import Box
from .grid import Grid
from .materials import (MAX_POLES, POLE_SHIFT, DebyePole, LorentzPole, MaterialMap,
                        tangential_edge_bits)
from .sources import GaussianPulse, port_sigma, sample_times

F32 = np.float32
PEC_SIGMA_THRESHOLD = 1e6  # rfx/api/_compile.py:118


@dataclass
class _Material:
    eps_r: float = 1.0
    sigma: float = 0.0
    mu_r: float = 1.0
    debye: tuple = ()          # DebyePole entries
    lorentz: tuple = ()        # LorentzPole entries


@dataclass
class _PortEntry:
    position: tuple
    component: str
    impedance: float
    waveform: GaussianPulse | None
    extent: float | None = None
    excite: bool = True


@dataclass
class _ProbeEntry:
    position: tuple
    component: str


@dataclass
class _DFTPlaneEntry:
    name: str
    axis: str
    coordinate: float
    component: str
    freqs: np.ndarray | None
    n_freqs: int


@dataclass
class _FluxMonitorEntry:
    name: str
    axis: str
    coordinate: float
    freqs: np.ndarray | None
    n_freqs: int
    size: tuple | None
    center: tuple | None
    window: str
    alpha: float


FLUX_COMPONENTS = {0: ("ey", "ez", "hy", "hz"), 1: ("ez", "ex", "hz", "hx"),
                   2: ("ex", "ey", "hx", "hy")}


@dataclass
class FluxMonitor:
    """Frequency-domain tangential E and H on a plane region
    (rfx/probes/probes.py:476-503); ``flux_spectrum`` integrates the flux."""
    e1_dft: np.ndarray    # complex128 (n_freqs, n1, n2)
    e2_dft: np.ndarray
    h1_dft: np.ndarray
    h2_dft: np.ndarray
    freqs: np.ndarray
    axis: int
    index: int
    dA: np.ndarray        # float32 (n1, n2) cell areas
    lo1: int
    hi1: int
    lo2: int
    hi2: int


def flux_spectrum(mon: FluxMonitor) -> np.ndarray:
    """Integral of Re(E x H*) . n dA per frequency, in float64
    (rfx/probes/probes.py:615-660); positive = power along +axis.  A monitor
    of differentiable torch tensors gives a torch tensor."""
    if isinstance(mon.e1_dft, torch.Tensor):
        dA = torch.as_tensor(np.asarray(mon.dA, np.float64), device=mon.e1_dft.device)
        integrand = (mon.e1_dft * torch.conj(mon.h2_dft)
                     - mon.e2_dft * torch.conj(mon.h1_dft))
        return torch.real(torch.sum(integrand * dA, dim=(-2, -1)))
    integrand = mon.e1_dft * np.conj(mon.h2_dft) - mon.e2_dft * np.conj(mon.h1_dft)
    return np.real(np.sum(integrand * np.asarray(mon.dA, np.float64), axis=(-2, -1)))


def dft_window_weights(total_steps: int, window: str = "rect", alpha: float = 0.25):
    """Streaming DFT window weight per step value 0..total_steps, float32
    (rfx/core/dft_utils.py:8-49); None for the rectangular window."""
    if window not in ("rect", "hann", "tukey"):
        raise ValueError(f"dft_window must be 'rect', 'hann', or 'tukey', got {window!r}")
    if total_steps <= 1 or window == "rect" or (window == "tukey" and alpha <= 0.0):
        return None
    x = np.arange(total_steps + 1, dtype=F32) / np.maximum(F32(total_steps - 1), F32(1.0))
    tp = F32(2.0 * np.pi)
    if window == "hann" or alpha >= 1.0:
        return (F32(0.5) * (F32(1.0) - np.cos(tp * x))).astype(F32)
    a = F32(alpha)
    left = F32(0.5) * (F32(1.0) + np.cos(F32(np.pi) * (F32(2.0) * x / a - F32(1.0))))
    right = F32(0.5) * (F32(1.0) + np.cos(F32(np.pi) * (F32(2.0) * x / a - F32(2.0) / a + F32(1.0))))
    return np.where(x < a / F32(2.0), left,
                    np.where(x <= F32(1.0) - a / F32(2.0), F32(1.0), right)).astype(F32)


class FDTDState(NamedTuple):
    ex: torch.Tensor
    ey: torch.Tensor
    ez: torch.Tensor
    hx: torch.Tensor
    hy: torch.Tensor
    hz: torch.Tensor
    step: int


@dataclass
class DFTPlaneProbe:
    """Frequency-domain plane accumulator (rfx/probes/probes.py:379-473)."""
    axis: int
    index: int
    component: str
    freqs: np.ndarray
    accumulator: np.ndarray   # complex128 (n_freqs, n1, n2); a torch tensor when differentiated


@dataclass
class Result:
    """Run output (rfx/api/_spec.py:649-692 subset)."""
    time_series: np.ndarray
    state: FDTDState | None = None
    dft_planes: dict = field(default_factory=dict)
    s_params: np.ndarray | None = None
    freqs: np.ndarray | None = None
    grid: object = None
    dt: float = 0.0
    flux_monitors: dict = field(default_factory=dict)
    ntff_data: object = None      # rfx_amd.farfield.NTFFData
    ntff_box: object = None       # rfx_amd.farfield.NTFFBox


@dataclass
class ForwardResult:
    """Differentiable forward output (rfx/api/_spec.py:997)."""
    time_series: torch.Tensor
    grid: object = None
    s_params: np.ndarray | None = None
    freqs: np.ndarray | None = None
    dft_planes: dict | None = None
    flux_monitors: dict | None = None
    ntff_data: object = None
    ntff_box: object = None


def _cell_cb_f32(dt, eps_r, sigma):
    """Cb at a cell in float32 (rfx/simulation.py:181-194)."""
    eps = F32(eps_r) * F32(EPS_0)
    loss = F32(sigma) * F32(dt) / (F32(2.0) * eps)
    return (F32(dt) / eps) / (F32(1.0) + loss)


def port_dft(samples: np.ndarray, freqs, dt: float, t_offset: int = 0):
    """sum_n x[n] exp(-j 2 pi f (n + t_offset) dt) dt in float64."""
    n = np.arange(samples.shape[0], dtype=np.float64) + t_offset
    ph = np.exp(-1j * 2.0 * np.pi * np.outer(np.asarray(freqs, np.float64), n * dt)) * dt
    return ph @ np.asarray(samples, np.float64)


def port_dft_torch(samples, freqs, dt: float, t_offset: int = 0):
    """port_dft on a torch tensor (differentiable, complex128)."""
    x = samples.to(torch.float64)
    n = torch.arange(x.shape[0], dtype=torch.float64, device=x.device) + t_offset
    f = torch.as_tensor(np.asarray(freqs, np.float64), device=x.device)
    ph = torch.exp(-2j * np.pi * torch.outer(f, n * dt)) * dt
    return ph @ x.to(torch.complex128)


def extract_lumped_s11_torch(v_dft, i_dft, z0=50.0):
    """extract_lumped_s11 on torch tensors."""
    den = v_dft - z0 * i_dft
    den = torch.where(den.abs() > 0.0, den, torch.ones_like(den))
    return (v_dft + z0 * i_dft) / den


def extract_lumped_s11(v_dft, i_dft, z0=50.0):
    """S11 = (V + Z0 I)/(V - Z0 I) (rfx/probes/probes.py:1239-1273)."""
    den = v_dft - z0 * i_dft
    den = np.where(np.abs(den) > 0.0, den, 1.0)
    return (v_dft + z0 * i_dft) / den


class Simulation:
    """Builder for one FDTD problem (rfx/api/__init__.py:280)."""

    def __init__(self, freq_max: float, domain, *, boundary: str | BoundarySpec = "cpml",
                 cpml_layers: int = 16, cpml_kappa_max: float = 1.0, pec_faces=None,
                 dx: float | None = None, mode: str = "3d", dz_profile=None,
                 dx_profile=None, dy_profile=None, precision: str = "float32",
                 solver: str = "yee", stencil_order: int = 2, device=None):
        if isinstance(boundary, BoundarySpec):
            # one kind on all six faces (rfx/api/__init__.py takes str or BoundarySpec)
            boundary, cpml_layers = run_kind(boundary, int(cpml_layers))
        if boundary not in ("pec", "cpml"):
            raise NotImplementedError(f"boundary={boundary!r} is not supported by the HIP backend "
                                      "(supported: 'pec', 'cpml')")
        if freq_max <= 0:
            raise ValueError(f"freq_max must be positive, got {freq_max}")
        if precision != "float32":
            raise NotImplementedError("the HIP backend computes in float32 only")
        if solver != "yee" or stencil_order != 2 or mode != "3d":
            raise NotImplementedError("only the 3-D second-order Yee solver is supported")
        if pec_faces:
            raise NotImplementedError("per-face PEC boundaries are not supported")
        self._freq_max = float(freq_max)
        # profiles set the domain extents of their axes; dz only where the
        # domain gives no z extent (rfx/api/__init__.py:434-476)
        prof = [None if q is None else np.asarray(q, np.float64)
                for q in (dx_profile, dy_profile, dz_profile)]
        dom = list(domain) + [0.0] * (3 - len(domain))
        if prof[0] is not None:
            dom[0] = float(np.sum(prof[0]))
        if prof[1] is not None:
            dom[1] = float(np.sum(prof[1]))
        if prof[2] is not None and dom[2] <= 0:
            dom[2] = float(np.sum(prof[2]))
        if (prof[0] is not None or prof[1] is not None) and dx is None:
            raise ValueError("dx_profile / dy_profile require an explicit dx (boundary cell size)")
        if any(d <= 0 for d in dom):
            raise ValueError(f"domain dimensions must be positive, got {tuple(dom)}")
        self._dx_profile, self._dy_profile, self._dz_profile = prof
        self._graded = any(q is not None for q in prof)
        self._domain = tuple(float(d) for d in dom)
        self._boundary = boundary
        self._cpml_layers = int(cpml_layers) if boundary == "cpml" else 0
        self._kappa_max = float(cpml_kappa_max)
        self._dx = dx
        self._materials: dict[str, _Material] = {}
        self._geometry: list[tuple[Box, str]] = []
        self._ports: list[_PortEntry] = []
        self._probes: list[_ProbeEntry] = []
        self._dft_planes: list[_DFTPlaneEntry] = []
        self._flux_monitors: list[_FluxMonitorEntry] = []
        self._ntff = None
        self._device = device

    # ---- builder --------------------------------------------------------------
    def add_material(self, name: str, *, eps_r: float = 1.0, sigma: float = 0.0,
                     mu_r: float = 1.0, debye_poles=None, lorentz_poles=None,
                     chi3: float = 0.0):
        """Material with optional dispersion (rfx/api/__init__.py:812-853):
        ``debye_poles`` a list of DebyePole(delta_eps, tau), ``lorentz_poles``
        a list of LorentzPole(omega_0, delta, kappa) (see ``lorentz_pole`` /
        ``drude_pole``); eps_r is then eps_inf.  Kerr chi3 is not
        supported."""
        if chi3:
            raise NotImplementedError("Kerr nonlinearity is not supported")
        poles = tuple(DebyePole(float(p[0]), float(p[1])) for p in (debye_poles or ()))
        for p in poles:
            if not p.tau > 0:
                raise ValueError(f"Debye pole {p}: tau must be positive")
        lor = tuple(LorentzPole(float(p[0]), float(p[1]), float(p[2]))
                    for p in (lorentz_poles or ()))
        self._materials[name] = _Material(float(eps_r), float(sigma), float(mu_r), poles, lor)
        return self

    def add(self, shape, *, material: str):
        if not isinstance(shape, Box):
            raise NotImplementedError("only Box shapes are supported")
        if material not in self._materials:
            raise KeyError(f"Unknown material {material!r}; register it with add_material()")
        self._geometry.append((shape, material))
        return self

    def add_source(self, position, component: str = "ez", *, waveform=None,
                   amplitude_kind: str | None = None):
        if component not in ("ex", "ey", "ez"):
            raise ValueError(f"component must be ex/ey/ez, got {component!r}")
        if amplitude_kind is not None:
            raise NotImplementedError("amplitude_kind is not supported (legacy meaning only)")
        if waveform is None:
            waveform = GaussianPulse(f0=self._freq_max / 2, bandwidth=0.8)
        self._ports.append(_PortEntry(tuple(position), component, 0.0, waveform))
        return self

    def add_port(self, position, component: str = "ez", *, impedance: float = 50.0,
                 waveform=None, extent: float | None = None, excite: bool = True,
                 direction=None, reference_plane_cells=None):
        if component not in ("ex", "ey", "ez"):
            raise ValueError(f"component must be ex/ey/ez, got {component!r}")
        if reference_plane_cells is not None:
            raise NotImplementedError("reference-plane port waves are not supported")
        if waveform is None and excite:
            waveform = GaussianPulse(f0=self._freq_max / 2, bandwidth=0.8)
        self._ports.append(_PortEntry(tuple(position), component, float(impedance), waveform,
                                      extent, excite))
        return self

    def add_probe(self, position, component: str = "ez"):
        if component not in ("ex", "ey", "ez", "hx", "hy", "hz"):
            raise ValueError(f"component must be a field name, got {component!r}")
        self._probes.append(_ProbeEntry(tuple(position), component))
        return self

    def add_vector_probe(self, position):
        for c in ("ex", "ey", "ez", "hx", "hy", "hz"):
            self.add_probe(position, c)
        return self

    def add_dft_plane_probe(self, *, axis: str, coordinate: float, component: str = "ez",
                            freqs=None, n_freqs: int = 50, name: str | None = None):
        if axis not in ("x", "y", "z"):
            raise ValueError(f"axis must be 'x', 'y', or 'z', got {axis!r}")
        if component not in ("ex", "ey", "ez", "hx", "hy", "hz"):
            raise ValueError(f"component must be a field name, got {component!r}")
        a = "xyz".index(axis)
        if coordinate < 0 or coordinate > self._domain[a]:
            raise ValueError(f"coordinate {coordinate} m is outside the {axis}-domain")
        if name is None:
            name = f"{component}_{axis}_{len(self._dft_planes)}"
        self._dft_planes.append(_DFTPlaneEntry(name, axis, float(coordinate), component,
                                               None if freqs is None else np.asarray(freqs),
                                               int(n_freqs)))
        return self

    def add_flux_monitor(self, *, axis: str, coordinate: float, freqs=None, n_freqs: int = 50,
                         size=None, center=None, name: str | None = None,
                         dft_window: str = "rect", dft_window_alpha: float = 0.25):
        """Poynting flux monitor on a plane (rfx/api/__init__.py:2577-2645):
        ``size`` / ``center`` restrict it to a region of the two tangential
        axes (uniform lane only, as in the reference's graded runner)."""
        if axis not in ("x", "y", "z"):
            raise ValueError(f"axis must be 'x', 'y', or 'z', got {axis!r}")
        a = "xyz".index(axis)
        if coordinate < 0 or coordinate > self._domain[a]:
            raise ValueError(f"coordinate {coordinate} m is outside the {axis}-domain "
                             f"[0, {self._domain[a]}]")
        if dft_window not in ("rect", "hann", "tukey"):
            raise ValueError(f"dft_window must be 'rect', 'hann', or 'tukey', got {dft_window!r}")
        if size is not None and self._graded:
            raise NotImplementedError("finite flux regions (size=...) are not supported on the "
                                      "graded mesh; use a full-plane monitor")
        if name is None:
            name = f"flux_{axis}_{len(self._flux_monitors)}"
        self._flux_monitors.append(_FluxMonitorEntry(
            name, axis, float(coordinate), None if freqs is None else np.asarray(freqs),
            int(n_freqs), None if size is None else tuple(size),
            None if center is None else tuple(center), dft_window, float(dft_window_alpha)))
        return self

    def add_ntff_box(self, corner_lo, corner_hi, freqs=None, n_freqs: int = 50):
        """Near-to-far-field box (rfx/api/__init__.py:2647-2669): the time
        loop accumulates the tangential-field DFTs on its six faces, returned
        as ``result.ntff_data`` / ``result.ntff_box`` for
        ``rfx_amd.farfield.compute_far_field``."""
        if freqs is None:
            freqs = np.linspace(self._freq_max / 10, self._freq_max, n_freqs)
        self._ntff = (tuple(corner_lo), tuple(corner_hi), np.asarray(freqs, F32))
        return self

    def _ntff_setup(self, grid):
        """NTFFBox for the registered box (rfx/runners/uniform.py:648-651,
        rfx/runners/nonuniform.py:1079-1097): corners through the lane's
        position-to-index rule, frequencies in float32."""
        if self._ntff is None:
            return None
        from .farfield import check_box, make_ntff_box
        lo, hi, freqs = self._ntff
        box = make_ntff_box(grid, lo, hi, freqs)
        check_box(box, grid.shape)
        return box

    def _ntff_results(self, box, dr, faces=None):
        if box is None:
            return None
        from .farfield import NTFFData, data_from_device
        if faces is not None:     # differentiable (n_freqs, n1, n2, 4) views
            return NTFFData(*[a.permute(1, 2, 3, 0) for a in faces])
        return data_from_device(dr.ntff_acc[0])

    def _flux_setup(self, grid, n_steps):
        """(Problem.fluxes entries, dA arrays) for the registered monitors
        (rfx/runners/uniform.py:66-137; rfx/runners/nonuniform.py:924-958)."""
        graded = self._graded
        out, areas = [], []
        for e in self._flux_monitors:
            a = "xyz".index(e.axis)
            pos = [0.0, 0.0, 0.0]
            pos[a] = e.coordinate
            index = grid.position_to_index(tuple(pos))[a]
            freqs = e.freqs if e.freqs is not None else np.linspace(
                self._freq_max / 10, self._freq_max, e.n_freqs).astype(F32)
            tang = [t for t in range(3) if t != a]
            if e.size is None:
                (lo1, hi1), (lo2, hi2) = (0, grid.shape[tang[0]]), (0, grid.shape[tang[1]])
            else:
                b = []
                for q, t in enumerate(tang):
                    c = e.center[q] if e.center is not None else self._domain[t] / 2.0
                    pad = grid.axis_pads[t]
                    lo = max(0, int(round(c / grid.dx - e.size[q] / (2.0 * grid.dx))) + pad)
                    hi = min(grid.shape[t], int(round(c / grid.dx + e.size[q] / (2.0 * grid.dx)))
                             + pad)
                    b.append((lo, hi))
                (lo1, hi1), (lo2, hi2) = b
            if hi1 <= lo1 or hi2 <= lo2:
                raise ValueError(f"flux monitor {e.name!r}: empty region")
            if graded:
                w1, w2 = grid.d32[tang[0]][lo1:hi1], grid.d32[tang[1]][lo2:hi2]
            else:
                w1 = np.full(hi1 - lo1, F32(grid.dx))
                w2 = np.full(hi2 - lo2, F32(grid.dx))
            areas.append(w1.reshape(-1, 1) * w2.reshape(1, -1))
            out.append((a, index, lo1, hi1, lo2, hi2, np.asarray(freqs, np.float64),
                        dft_window_weights(n_steps, e.window, e.alpha)))
        return out, areas

    def _flux_results(self, prob, dr, areas, accs=None):
        res = {}
        for e, (a, idx, lo1, hi1, lo2, hi2, fr, _), acc, dA in zip(
                self._flux_monitors, prob.fluxes, dr.flux_acc if accs is None else accs, areas):
            x = acc.cpu().numpy() if accs is None else acc
            res[e.name] = FluxMonitor(x[0], x[1], x[2], x[3], np.asarray(fr), a, int(idx), dA,
                                      lo1, hi1, lo2, hi2)
        return res

    # ---- set-up ---------------------------------------------------------------
    def _build_grid(self):
        if self._graded:
            from .nonuniform import make_nonuniform_grid
            return make_nonuniform_grid(self._freq_max, self._domain, self._dz_profile,
                                        self._dx, self._cpml_layers, self._dx_profile,
                                        self._dy_profile)
        return Grid(self._freq_max, self._domain, self._dx, self._cpml_layers)

    def _build_problem(self, n_steps: int, lane: str, eps_override=None,
                       sigma_override=None, mu_r_override=None, s11_freqs=None,
                       material_sweep=None, subpixel_smoothing=False, pec_mask_override=None):
        """Host set-up of one run.  ``material_sweep = (name, field, value)``
        replaces one material property as rfx/vmap_sweep.py:265-390 does:
        with a material name, the cells inside that material's shapes take
        the value before the CPML pads are extended; without one, every
        material whose property differs from vacuum takes it, pads included."""
        grid = self._build_grid()
        graded = self._graded
        shape = grid.shape
        mat = MaterialMap(shape)
        coords = grid.node_coords()
        pec = None
        has_poles = any(self._materials[name].debye or self._materials[name].lorentz
                        for _, name in self._geometry)
        overridden = (eps_override is not None or sigma_override is not None
                      or mu_r_override is not None or pec_mask_override is not None)
        dasm = None
        if self._use_device_assembly(shape, material_sweep is not None or has_poles or overridden):
            # box fill, pads and PEC bits on the GPU; single-cell edits (ports)
            # go through the same interface as on the host map
            from .assembly import DeviceAssembly
            dasm = DeviceAssembly(mat, self._device or "cuda")
            boxes = []
            for box, name in self._geometry:
                m = self._materials[name]
                props = None if m.sigma >= PEC_SIGMA_THRESHOLD else (m.eps_r, m.sigma, m.mu_r)
                boxes.append((box.axis_masks(coords), props))
            dasm.fill(boxes, grid.face_pads if self._cpml_layers > 0 else None)
            pec = dasm.pec
            mat = dasm
        for box, name in ([] if dasm is not None else self._geometry):
            m = self._materials[name]
            masks = box.axis_masks(coords)
            if m.sigma >= PEC_SIGMA_THRESHOLD:
                if pec is None:
                    pec = np.zeros(shape, bool)
                pec[np.ix_(*[np.nonzero(x)[0] for x in masks])] = True
            else:
                mat.fill_box(masks, m.eps_r, m.sigma, m.mu_r)
        if pec_mask_override is not None:
            # merged with the geometry's PEC cells (rfx/api/_execute.py:2807-2808)
            extra = np.asarray(pec_mask_override, bool)
            if extra.shape != tuple(shape):
                raise ValueError(f"pec_mask_override shape {extra.shape} != grid shape {shape}")
            pec = extra.copy() if pec is None else (pec | extra)
        if material_sweep is not None and material_sweep[0] is not None:
            sw_name, sw_field, sw_value = material_sweep
            sel = np.zeros(shape, bool)
            for box, name in self._geometry:
                if name == sw_name:
                    sel[np.ix_(*[np.nonzero(x)[0] for x in box.axis_masks(coords)])] = True
            mat.set_field_where(sel, sw_field, sw_value)
        if self._cpml_layers > 0 and dasm is None:
            mat.extend_pads(grid.face_pads)
        if material_sweep is not None and material_sweep[0] is None:
            mat.set_field_all(material_sweep[1], material_sweep[2])
        # Debye pole masks (rfx/api/_compile.py:203-232, 352): per pole, keyed by
        # its value in order of first appearance, OR-ed over every shape of a
        # material listing it (a later shape does not remove it), not extended
        # into the CPML pads; carried as flag bits 16 + q of the material ids
        # (Debye poles take bits 16.., Lorentz poles the bits after them)
        poles, lorentz = [], []
        for box, name in self._geometry:
            m = self._materials[name]
            poles += [p for p in m.debye if p not in poles]
            lorentz += [p for p in m.lorentz if p not in lorentz]
        if len(poles) + len(lorentz) > MAX_POLES:
            raise NotImplementedError(f"more than {MAX_POLES} distinct dispersion poles")
        if poles or lorentz:
            pole_bits = np.zeros(shape, np.uint32)
            for box, name in self._geometry:
                m = self._materials[name]
                bits = sum(1 << (POLE_SHIFT + poles.index(p)) for p in set(m.debye))
                bits |= sum(1 << (POLE_SHIFT + len(poles) + lorentz.index(p))
                            for p in set(m.lorentz))
                if bits:
                    sel = np.ix_(*[np.nonzero(x)[0] for x in box.axis_masks(coords)])
                    pole_bits[sel] |= np.uint32(bits)
            mat.add_flags(pole_bits)
        # graded lane: sources are normalised by the materials before any
        # override or port fold (rfx/runners/nonuniform.py:630-643)
        concrete = {}
        if graded:
            for pe in self._ports:
                for c in self._port_cells(grid, pe):
                    concrete[c] = mat.cell_props(c)[:2]
        if (eps_override is not None or sigma_override is not None
                or mu_r_override is not None) and (lane == "forward" or graded):
            mat.override(eps_override, sigma_override, mu_r_override)
        dt = grid.dt
        times = sample_times(n_steps, dt)
        sources, ports = [], []
        for pe in self._ports:
            if graded:
                self._graded_port(grid, mat, pe, times, dt, concrete, pec, sources, ports,
                                  s11_freqs)
                continue
            idx = grid.position_to_index(pe.position)
            if pe.impedance == 0.0:
                wave = pe.waveform(times)
                if self._boundary == "cpml" or lane == "forward":
                    e, s, _, _ = mat.cell_props(idx)
                    wave = (_cell_cb_f32(dt, e, s) * wave).astype(F32)
                sources.append((idx, pe.component, wave))
                continue
            if pe.extent is not None:
                raise NotImplementedError("wire ports (extent=...) are supported on the graded "
                                          "lane only")
            sizes = (grid.dx, grid.dx, grid.dx)
            sp = port_sigma(sizes, pe.component, pe.impedance)
            e, s, mu, fl = mat.cell_props(idx)
            mat.set_cell(idx, sig=F32(s + F32(sp)))
            if pec is not None:
                pec[idx] = False
            if pe.excite:
                e, s, _, _ = mat.cell_props(idx)
                cb = _cell_cb_f32(dt, e, s)
                wave = ((cb / F32(grid.dx)) * pe.waveform(times)).astype(F32)
                sources.append((idx, pe.component, wave))
            if s11_freqs is not None:
                ports.append((idx, pe.component, float(grid.dx), (grid.dx,) * 3, pe.impedance,
                              pe.excite))
        if dasm is not None:
            mat = dasm.finish()
        elif pec is not None:
            mat.add_flags(tangential_edge_bits(pec))
        if subpixel_smoothing:
            self._apply_subpixel(grid, mat, subpixel_smoothing, bool(poles or lorentz))
        probes = [(grid.position_to_index(p.position), p.component) for p in self._probes]
        if not probes and lane == "forward" and self._ports:
            probes = [(grid.position_to_index(p.position), p.component) for p in self._ports]
        planes = []
        for pe in self._dft_planes:
            a = "xyz".index(pe.axis)
            pos = [0.0, 0.0, 0.0]
            pos[a] = pe.coordinate
            index = grid.position_to_index(tuple(pos))[a]
            freqs = pe.freqs if pe.freqs is not None else np.linspace(
                self._freq_max / 10, self._freq_max, pe.n_freqs).astype(F32)
            planes.append((a, index, pe.component, np.asarray(freqs, np.float64)))
        inv_e, inv_h = grid.inverse_spacings()
        # the graded grid carries no kappa_max, so its CPML uses 1
        # (rfx/boundaries/cpml.py:391-392)
        kappa = 1.0 if graded else self._kappa_max
        faces = face_tables(self._cpml_layers, dt, grid.face_cell_sizes(), kappa,
                            "xyz" if self._cpml_layers > 0 else "")
        # run(): outer-face PEC on every axis (runners/uniform.py:321 passes
        # pec_axes=None, simulation.py:684-691).  forward(): pec_axes = axes
        # outside grid.cpml_axes (_execute.py:1420-1421); Grid keeps
        # cpml_axes="xyz" even with zero layers, so forward() applies none.
        # The graded lane always applies it (nonuniform.py:1319).
        pec_axes = "xyz" if (graded or lane == "run") else ""
        fluxes, areas = self._flux_setup(grid, n_steps)
        box = self._ntff_setup(grid)
        ntff = [] if box is None else [(box.lo, box.hi, np.asarray(box.freqs, np.float64))]
        prob = Problem(shape=shape, dt=dt, graded=graded, inv_e=inv_e, inv_h=inv_h, mat=mat,
                       faces=faces, pec_axes=pec_axes, sources=sources, probes=probes,
                       ports=[p[:4] for p in ports], planes=planes, fluxes=fluxes,
                       record_order="graded" if graded else "uniform", ntff=ntff,
                       poles=[(p.delta_eps, p.tau) for p in poles],
                       lorentz=[(p.omega_0, p.delta, p.kappa) for p in lorentz])
        prob.flux_areas = areas
        prob.ntff_box = box
        return grid, prob, ports

    def _apply_subpixel(self, grid, mat, how, dispersive: bool):
        """Per-component eps at the interface cells (rfx/runners/uniform.py:
        234-278, graded rfx/runners/nonuniform.py:648-672): every geometry
        entry with its material's eps_r, background 1.  Dispersive problems
        keep the plain update (the reference's dispersive E update does not
        read the smoothed eps)."""
        if how == "kottke_pec":
            raise NotImplementedError("subpixel_smoothing='kottke_pec' (the inverse-permittivity "
                                      "PEC tensor) is not supported by the HIP backend")
        if how is not True:
            raise ValueError(f"subpixel_smoothing must be True, False or 'kottke_pec', got {how!r}")
        if dispersive:
            warnings.warn("subpixel_smoothing=True has no effect with dispersive (Debye / "
                          "Lorentz) materials", stacklevel=3)
            return
        from .smoothing import smoothed_eps
        boxes = [(box.corner_lo, box.corner_hi, self._materials[name].eps_r)
                 for box, name in self._geometry]
        if boxes:
            mat.apply_aniso(*smoothed_eps(grid, boxes))

    # material assembly on the GPU: "auto" (grids of DEVICE_ASSEMBLY_CELLS cells
    # or more, when a GPU is visible), "device" or "host"
    material_assembly = "auto"
    DEVICE_ASSEMBLY_CELLS = 1 << 21

    def _use_device_assembly(self, shape, host_only: bool) -> bool:
        """Device assembly covers box fill, pads, port cells and PEC bits;
        material sweeps, dispersion poles and overrides edit the ids on the
        host, so those problems keep the host assembly."""
        how = self.material_assembly
        if how not in ("auto", "device", "host"):
            raise ValueError(f"material_assembly must be 'auto', 'device' or 'host', got {how!r}")
        if how == "host" or host_only:
            if how == "device" and host_only:
                raise NotImplementedError("device material assembly does not cover material "
                                          "sweeps, dispersion poles or overrides")
            return False
        if not torch.cuda.is_available():
            if how == "device":
                from ._lib import RfxHipError
                raise RfxHipError("device material assembly needs a GPU")
            return False
        return how == "device" or int(np.prod(shape)) >= self.DEVICE_ASSEMBLY_CELLS

    def _port_cells(self, grid, pe):
        """Cells a port entry touches (wire ports: every extent cell)."""
        idx = grid.position_to_index(pe.position)
        if pe.impedance == 0.0 or pe.extent is None:
            return [idx]
        axis = "xyz".index(pe.component[1])
        end = list(pe.position)
        end[axis] += pe.extent
        k_end = grid.position_to_index(tuple(end))[axis]
        lo, hi = sorted((idx[axis], k_end))
        cells = []
        for k in range(lo, hi + 1):
            c = list(idx)
            c[axis] = k
            cells.append(tuple(c))
        return cells

    def _graded_port(self, grid, mat, pe, times, dt, concrete, pec, sources, ports, s11_freqs):
        """One port entry on the graded lane (rfx/runners/nonuniform.py:
        694-887): soft source -> current source; single-cell port -> sigma
        fold d_par/(Z0 dual1 dual2) + current source; wire port -> fold
        n_live d/(Z0 dual1 dual2) on every live extent cell, source/n_live on
        each, V/I recorded at the middle cell."""
        from .nonuniform import current_source_scale
        axis = "xyz".index(pe.component[1])
        cells = self._port_cells(grid, pe)

        def source(c):
            e, s = concrete[c]
            return (current_source_scale(grid, e, s, c, pe.component, dt)
                    * pe.waveform(times)).astype(F32)

        if pe.impedance == 0.0:
            sources.append((cells[0], pe.component, source(cells[0])))
            return
        live = [pec is None or not bool(pec[c]) for c in cells]
        n_live = sum(live)
        if n_live == 0:
            raise ValueError(f"port at {pe.position}: every extent cell lies inside PEC "
                             "geometry, so it has no live cell to terminate or drive")
        wire = pe.extent is not None
        scale = n_live if wire else 1
        for c, lv in zip(cells, live):
            if not lv:
                continue
            d_par, p1, p2, _ = grid.port_metrics(c, axis)
            _, s, _, _ = mat.cell_props(c)
            mat.set_cell(c, sig=F32(s + F32(scale * d_par / (pe.impedance * p1 * p2))))
            if pec is not None:
                pec[c] = False
        if pe.excite:
            for c, lv in zip(cells, live):
                if lv:
                    w = source(c)
                    sources.append((c, pe.component, (w / n_live).astype(F32) if wire else w))
        if wire and s11_freqs is not None:
            mid = cells[len(cells) // 2]
            d_par, _, _, loop = grid.port_metrics(mid, axis)
            ports.append((mid, pe.component, d_par, loop, pe.impedance, pe.excite))

    # ---- run / forward -----------------------------------------------------------
    def run(self, n_steps: int | None = None, *, num_periods: float = 20.0,
            use_graph: bool = True, s_param_freqs=None, autotune: bool = False,
            subpixel_smoothing: bool | str = False, **unsupported):
        """Run the time loop (rfx/api/_execute.py:2862 -> runners/uniform.py:140,
        graded z -> runners/nonuniform.py:465).  On the graded lane
        ``s_param_freqs`` gives the wire-port S matrix S[j, k] = b_j / a_k
        (rfx/nonuniform.py:1900-1935).  ``autotune=True`` first picks the
        fastest kernel variant for this problem (a few timed steps, then a
        restart from zero).  ``subpixel_smoothing=True`` gives the cells at
        material interfaces per-component Kottke-averaged eps
        (rfx/geometry/smoothing.py; those plans run the two-pass kernels)."""
        _reject(unsupported, "run")
        graded = self._graded
        grid = self._build_grid()
        if n_steps is None:
            n_steps = grid.num_timesteps(num_periods)
        if not graded and any(p.impedance != 0.0 for p in self._ports):
            warnings.warn("run(): lumped-port S-parameters come from forward("
                          "port_s11_freqs=...); run() returns no s_params", stacklevel=2)
        sf = None
        if graded and s_param_freqs is not None:
            sf = np.asarray(s_param_freqs, F32)
        grid, prob, ports = self._build_problem(int(n_steps), "run", s11_freqs=sf,
                                                subpixel_smoothing=subpixel_smoothing)
        dr = DeviceRun(prob, int(n_steps), self._device)
        if autotune:
            # time the kernel variants on this problem's first steps, keep
            # the fastest, restart from a zero state (rfx_amd/autotune.py)
            from .autotune import autotune as _tune
            _tune(dr, trial_steps=min(32, int(n_steps)))
        dr.run(0, int(n_steps), use_graph=use_graph)
        torch.cuda.synchronize(dr.device)
        ts = dr.time_series().cpu().numpy().astype(np.float32)
        s_params = None
        if ports:
            samp = dr.port_samples().cpu().numpy().astype(np.float64)
            ab = []
            for q, p in enumerate(ports):
                v = port_dft(samp[:, q, 0], sf, grid.dt)
                i = port_dft(samp[:, q, 1], sf, grid.dt)
                zi = p[4] * i
                ab.append(((-v + zi) / 2.0, (-v - zi) / 2.0))
            excited = [q for q, p in enumerate(ports) if p[5]] or list(range(len(ports)))
            s_params = np.zeros((len(ports), len(ports), len(sf)), np.complex128)
            for k in excited:
                a_k = np.where(np.abs(ab[k][0]) > 0, ab[k][0], 1.0)
                for j in range(len(ports)):
                    s_params[j, k] = ab[j][1] / a_k
        planes = {}
        for entry, (a, idx, comp, freqs), acc in zip(self._dft_planes, prob.planes, dr.plane_acc):
            planes[entry.name] = DFTPlaneProbe(a, int(idx), comp, np.asarray(freqs),
                                               acc.cpu().numpy())
        st = FDTDState(*[f for f in dr.fields], step=int(n_steps))
        _warn_if_nonfinite(ts, "run")
        return Result(time_series=ts, state=st, dft_planes=planes, s_params=s_params,
                      freqs=sf if s_params is not None else None, grid=grid, dt=grid.dt,
                      flux_monitors=self._flux_results(prob, dr, prob.flux_areas),
                      ntff_data=self._ntff_results(prob.ntff_box, dr), ntff_box=prob.ntff_box)

    def forward(self, *, eps_override=None, sigma_override=None, mu_r_override=None,
                pec_mask_override=None, n_steps: int | None = None, num_periods: float = 20.0,
                checkpoint: bool = True, checkpoint_segments: int | None = None,
                port_s11_freqs=None, skip_preflight: bool = False, grad_region=None,
                emit_time_series: bool = True, checkpoint_every: int | None = None,
                n_warmup: int = 0, **unsupported):
        """Differentiable forward (rfx/api/_execute.py:2447-2858): no outer-face
        PEC on axes covered by CPML (_execute.py:1420-1421); the returned
        time series (and, with lumped ports and ``port_s11_freqs``, the
        S-parameters, a torch tensor then) carry gradients w.r.t.
        ``eps_override``, ``sigma_override`` and ``mu_r_override`` through the
        hand-written adjoint.  ``pec_mask_override`` (bool, grid shape) adds
        hard PEC cells to the geometry's (rfx/api/_execute.py:2807-2808).
        ``grad_region`` = (i0, i1, j0, j1, k0, k1)
        limits where dJ/d eps_r is evaluated (default: the whole grid);
        ``checkpoint_segments`` = K forces K recompute segments.  On the
        graded lane, as in the reference (rfx/nonuniform.py:1589-1804):
        ``checkpoint_every`` = recompute segment length, ``n_warmup`` = first
        steps whose end state is a constant of the gradient (a truncated
        gradient; outputs unchanged), ``emit_time_series=False`` = no probe
        series; the uniform lane refuses them as the reference's does
        (rfx/api/_execute.py:2331-2357)."""
        _reject(unsupported, "forward")
        if not self._graded and (not emit_time_series or checkpoint_every is not None
                                 or n_warmup > 0):
            raise NotImplementedError(
                "emit_time_series=False, checkpoint_every and n_warmup are supported on the "
                "non-uniform (graded) forward path only; on the uniform lane use "
                "checkpoint_segments")
        grid = self._build_grid()
        if n_steps is None:
            n_steps = grid.num_timesteps(num_periods)
        n_steps = int(n_steps)
        n_warmup = int(n_warmup)
        if n_warmup and n_warmup >= n_steps:
            raise ValueError(f"n_warmup ({n_warmup}) must be < n_steps ({n_steps})")
        if checkpoint_every is not None and 0 < int(checkpoint_every) < n_steps - n_warmup:
            checkpoint_segments = math.ceil((n_steps - n_warmup) / int(checkpoint_every))
        s11 = None if port_s11_freqs is None else np.asarray(port_s11_freqs, F32)
        def _req(t):
            return isinstance(t, torch.Tensor) and t.requires_grad

        needs_grad = _req(eps_override) or _req(sigma_override) or _req(mu_r_override)
        eps_np = _to_np(eps_override)
        if isinstance(pec_mask_override, torch.Tensor):
            pec_mask_override = pec_mask_override.detach().cpu().numpy()
        grid, prob, ports = self._build_problem(n_steps, "forward", eps_np,
                                                _to_np(sigma_override), _to_np(mu_r_override),
                                                s11, pec_mask_override=pec_mask_override)
        port_t = plane_t = flux_t = None
        if needs_grad:
            from .adjoint import forward_with_grad
            ts, port_t, plane_t, flux_t, dr = forward_with_grad(self, prob, grid, eps_override, n_steps,
                                               checkpoint_segments, grad_region,
                                               sigma_tensor=sigma_override,
                                               mu_tensor=mu_r_override, n_warmup=n_warmup)
        else:
            dr = DeviceRun(prob, n_steps, self._device)
            dr.run(0, n_steps)
            ts = dr.time_series().clone()
        s_params = freqs = None
        if ports and port_t is not None:
            # differentiable S11: the same DFT and extraction in torch (complex128)
            s_list = [extract_lumped_s11_torch(port_dft_torch(port_t[:, q, 0], s11, grid.dt),
                                               port_dft_torch(port_t[:, q, 1], s11, grid.dt),
                                               p[4]) for q, p in enumerate(ports)]
            s_params = s_list[0] if len(s_list) == 1 else torch.stack(s_list)
            freqs = s11
        elif ports:
            torch.cuda.synchronize(dr.device)
            samp = dr.port_samples().cpu().numpy().astype(np.float64)
            s_list = []
            for q, p in enumerate(ports):
                v = port_dft(samp[:, q, 0], s11, grid.dt)
                i = port_dft(samp[:, q, 1], s11, grid.dt)
                s_list.append(extract_lumped_s11(v, i, p[4]))
            s_params = s_list[0] if len(s_list) == 1 else np.stack(s_list)
            freqs = s11
        planes = None
        if self._dft_planes:
            # with a gradient requested the accumulators are differentiable
            # torch tensors (complex128), else numpy arrays
            accs = plane_t if plane_t is not None else [a.cpu().numpy() for a in dr.plane_acc]
            planes = {e.name: DFTPlaneProbe(a, int(idx), comp, np.asarray(fr), acc)
                      for e, (a, idx, comp, fr), acc in zip(self._dft_planes, prob.planes, accs)}
        # with a gradient requested, flux monitors and NTFF faces hold
        # differentiable torch tensors
        n_mon = len(prob.fluxes)
        fluxes = (self._flux_results(prob, dr, prob.flux_areas,
                                     None if flux_t is None else flux_t[:n_mon])
                  if self._flux_monitors else None)
        ntff = self._ntff_results(prob.ntff_box, dr,
                                  None if flux_t is None else flux_t[n_mon:n_mon + 6])
        if not emit_time_series:      # no probe series (rfx/nonuniform.py:1538-1542)
            ts = ts[:, :0]
        return ForwardResult(time_series=ts, grid=grid, s_params=s_params, freqs=freqs,
                             dft_planes=planes, flux_monitors=fluxes,
                             ntff_data=ntff, ntff_box=prob.ntff_box)


def _to_np(x):
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        return x.detach().to("cpu", torch.float32).numpy()
    return np.asarray(x, F32)


def _reject(kw, where):
    if kw:
        raise NotImplementedError(f"{where}(): unsupported arguments {sorted(kw)}")


def _warn_if_nonfinite(ts, context):
    if ts.size and not np.all(np.isfinite(ts)):
        warnings.warn(f"{context}: non-finite values in the probe time series", stacklevel=3)
