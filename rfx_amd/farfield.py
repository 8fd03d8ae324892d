"""Near-to-far-field transform (Huygens box) on the HIP backend.

The time loop accumulates, on the six faces of a box, the DFTs of the
tangential E and H (include/rfxhip.h ``rfxh_ntff``; the reference's
rfx/farfield.py:208-313 ``accumulate_ntff``).  ``compute_far_field`` turns
them into the far-field components E_theta, E_phi with the surface
equivalence principle (rfx/farfield.py:393-556) in one HIP kernel
(``rfxh_ntff_far_field``).  The data types and the position / area rules
follow the reference so that results drop in:

* ``NTFFBox``: grid indices of the box faces and the frequencies (float32,
  as rfx/farfield.py:127-141 ``make_ntff_box`` stores them);
* ``NTFFData``: per face (n_freqs, n1, n2, 4) complex128, x faces
  [ey, ez, hy, hz], y faces [ex, ez, hx, hz], z faces [ex, ey, hx, hy];
* face positions (rfx/farfield.py:344-390): (index - cpml_lo) * dx on the
  uniform axes, cumulative z edges on a graded z axis; the lo offset is the
  box's own CPML thickness or else ``grid.cpml_layers``;
* face areas (rfx/farfield.py:456-462): dx * dy, or d_perp * dz[k] on the
  x / y faces of a graded z axis.  On the graded lane the x and y positions
  and widths use the boundary spacings grid.dx, grid.dy, as the reference's
  transform does.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .constants import C0, EPS_0, MU_0

ETA_0 = float(np.sqrt(MU_0 / EPS_0))   # rfx/farfield.py:25
F32 = np.float32


class NTFFBox(NamedTuple):
    """Huygens box by grid indices (rfx/farfield.py:32-78)."""
    i_lo: int
    i_hi: int
    j_lo: int
    j_hi: int
    k_lo: int
    k_hi: int
    freqs: np.ndarray
    cpml_lo_x: int = 0
    cpml_hi_x: int = 0
    cpml_lo_y: int = 0
    cpml_hi_y: int = 0
    cpml_lo_z: int = 0
    cpml_hi_z: int = 0

    @property
    def lo(self):
        return (self.i_lo, self.j_lo, self.k_lo)

    @property
    def hi(self):
        return (self.i_hi, self.j_hi, self.k_hi)


class NTFFData(NamedTuple):
    """Accumulated tangential-field DFTs per face, (n_freqs, n1, n2, 4)
    complex128 (rfx/farfield.py:81-104, float64 sums here)."""
    x_lo: np.ndarray
    x_hi: np.ndarray
    y_lo: np.ndarray
    y_hi: np.ndarray
    z_lo: np.ndarray
    z_hi: np.ndarray


class FarFieldResult(NamedTuple):
    """E_theta, E_phi: (n_freqs, n_theta, n_phi) complex128, without the
    1/r factor (rfx/farfield.py:107-124)."""
    E_theta: np.ndarray
    E_phi: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    freqs: np.ndarray


FACES = ("x_lo", "x_hi", "y_lo", "y_hi", "z_lo", "z_hi")


def make_ntff_box(grid, corner_lo, corner_hi, freqs) -> NTFFBox:
    """Box from physical corners (rfx/farfield.py:127-141)."""
    lo = grid.position_to_index(tuple(corner_lo))
    hi = grid.position_to_index(tuple(corner_hi))
    return NTFFBox(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                   np.asarray(freqs, F32))


def check_box(box: NTFFBox, shape):
    for a in range(3):
        if not 0 <= box.lo[a] < box.hi[a] < shape[a]:
            raise ValueError(
                f"NTFF box {'xyz'[a]} indices [{box.lo[a]}, {box.hi[a]}] must satisfy "
                f"0 <= lo < hi < {shape[a]}")


def face_shapes(box: NTFFBox):
    ni, nj, nk = (box.hi[a] - box.lo[a] for a in range(3))
    return [(nj, nk), (nj, nk), (ni, nk), (ni, nk), (ni, nj), (ni, nj)]


def data_from_device(accs) -> NTFFData:
    """Six device accumulators [4][n_freqs][n1][n2] -> NTFFData."""
    return NTFFData(*[a.cpu().numpy().transpose(1, 2, 3, 0).copy() for a in accs])


def _axis_tables(box: NTFFBox, grid):
    """Per-axis positions and widths (float64) by the reference's rules."""
    legacy = int(getattr(grid, "cpml_layers", 0) or 0)
    cpml_lo = (box.cpml_lo_x or legacy, box.cpml_lo_y or legacy, box.cpml_lo_z or legacy)
    dx = float(grid.dx)
    dy = float(getattr(grid, "dy", dx))
    shape = grid.shape
    pos, width = [], []
    for a, d in ((0, dx), (1, dy)):
        pos.append((np.arange(shape[a], dtype=np.float64) - cpml_lo[a]) * d)
        width.append(np.full(shape[a], d))
    dz = getattr(grid, "dz", None)
    if dz is not None:   # graded z: cumulative cell edges from the float32 profile
        dz64 = np.asarray(dz, np.float64)
        edges = np.concatenate([[0.0], np.cumsum(dz64)])
        edges = edges - edges[cpml_lo[2]]
        pos.append(np.ascontiguousarray(edges[:shape[2]]))
        width.append(np.ascontiguousarray(dz64))
    else:
        pos.append((np.arange(shape[2], dtype=np.float64) - cpml_lo[2]) * dx)
        width.append(np.full(shape[2], dx))
    return pos, width


def _far_field_desc(box: NTFFBox, grid, freqs):
    """The transform's descriptor without accumulators, and the host tables
    it points at (kept alive by the caller for the call)."""
    pos, width = _axis_tables(box, grid)
    d = _lib.FarFieldDesc()
    for a in range(3):
        d.lo[a], d.hi[a] = box.lo[a], box.hi[a]
        d.n_axis[a] = grid.shape[a]
        d.pos[a] = pos[a].ctypes.data_as(C.POINTER(C.c_double))
        d.width[a] = width[a].ctypes.data_as(C.POINTER(C.c_double))
    d.n_freq = len(freqs)
    d.freqs = freqs.ctypes.data_as(C.POINTER(C.c_double))
    d.c0, d.eta0 = C0, ETA_0
    return d, (pos, width, freqs)


def _run_far_field(box, grid, theta, phi, freqs, accs, dev):
    """E_theta, E_phi on the device from six device accumulators
    [4][n_freqs][n1][n2] (rfxh_ntff_far_field)."""
    lib = _lib.load()
    d, keep = _far_field_desc(box, grid, freqs)
    for q in range(6):
        d.d_acc[q] = accs[q].data_ptr()
    e_th = torch.empty((len(freqs), len(theta), len(phi)), dtype=torch.complex128, device=dev)
    e_ph = torch.empty_like(e_th)
    dp = C.POINTER(C.c_double)
    with torch.cuda.device(dev):
        _lib.check(lib.rfxh_ntff_far_field(
            C.byref(d), len(theta), theta.ctypes.data_as(dp), len(phi), phi.ctypes.data_as(dp),
            e_th.data_ptr(), e_ph.data_ptr(), torch.cuda.current_stream(dev).cuda_stream))
    del keep
    return e_th, e_ph


class _FarFieldFn(torch.autograd.Function):
    """The transform as a torch function of the six face accumulators: the
    forward is rfxh_ntff_far_field, the backward its conjugate transpose
    (rfxh_ntff_far_field_adjoint)."""

    @staticmethod
    def forward(ctx, setup, *accs):
        box, grid, theta, phi, freqs, dev = setup
        ctx.setup = setup
        ctx.shapes = [a.shape for a in accs]
        ctx.set_materialize_grads(False)
        return _run_far_field(box, grid, theta, phi, freqs, [a.contiguous() for a in accs], dev)

    @staticmethod
    def backward(ctx, g_th, g_ph):
        box, grid, theta, phi, freqs, dev = ctx.setup
        shape = (len(freqs), len(theta), len(phi))
        zero = torch.zeros(shape, dtype=torch.complex128, device=dev)
        g_th = zero if g_th is None else g_th.to(dev, torch.complex128).contiguous()
        g_ph = zero if g_ph is None else g_ph.to(dev, torch.complex128).contiguous()
        out = [torch.empty(sh, dtype=torch.complex128, device=dev) for sh in ctx.shapes]
        ptrs = (C.c_void_p * 6)(*[o.data_ptr() for o in out])
        lib = _lib.load()
        d, keep = _far_field_desc(box, grid, freqs)
        dp = C.POINTER(C.c_double)
        with torch.cuda.device(dev):
            _lib.check(lib.rfxh_ntff_far_field_adjoint(
                C.byref(d), len(theta), theta.ctypes.data_as(dp), len(phi),
                phi.ctypes.data_as(dp), g_th.data_ptr(), g_ph.data_ptr(),
                C.cast(ptrs, C.POINTER(C.c_void_p)), torch.cuda.current_stream(dev).cuda_stream))
        del keep
        return (None, *out)


def compute_far_field(ntff_data: NTFFData, box: NTFFBox, grid, theta, phi,
                      device=None) -> FarFieldResult:
    """E_theta, E_phi on the (theta, phi) grid (rfx/farfield.py:393-556),
    computed by the rfxh_ntff_far_field kernel in float64.  Face data held
    as torch tensors (``forward()`` with a gradient requested gives such
    data) give torch results on their device, differentiable through the
    transform's adjoint (rfxh_ntff_far_field_adjoint)."""
    _lib.require_device()
    theta = np.ascontiguousarray(np.asarray(theta, np.float64).ravel())
    phi = np.ascontiguousarray(np.asarray(phi, np.float64).ravel())
    freqs = np.ascontiguousarray(np.asarray(box.freqs, np.float64).ravel())
    check_box(box, grid.shape)
    nf = len(freqs)
    faces = [getattr(ntff_data, name) for name in FACES]
    for name, shp, x in zip(FACES, face_shapes(box), faces):
        if tuple(x.shape) != (nf,) + tuple(shp) + (4,):
            raise ValueError(f"NTFF face {name}: shape {tuple(x.shape)}, "
                             f"expected {(nf,) + tuple(shp) + (4,)}")
    if any(isinstance(x, torch.Tensor) for x in faces):
        dev = torch.device(device) if device is not None else next(
            x.device for x in faces if isinstance(x, torch.Tensor))
        accs = [torch.as_tensor(x).to(dev, torch.complex128).permute(3, 0, 1, 2) for x in faces]
        e_th, e_ph = _FarFieldFn.apply((box, grid, theta, phi, freqs, dev), *accs)
        return FarFieldResult(e_th, e_ph, theta, phi, freqs)
    dev = torch.device(device if device is not None else "cuda")
    accs = [torch.from_numpy(np.ascontiguousarray(
        np.asarray(x).astype(np.complex128).transpose(3, 0, 1, 2))).to(dev) for x in faces]
    e_th, e_ph = _run_far_field(box, grid, theta, phi, freqs, accs, dev)
    return FarFieldResult(e_th.cpu().numpy(), e_ph.cpu().numpy(), theta, phi, freqs)


def radiation_pattern(ff: FarFieldResult) -> np.ndarray:
    """Normalised pattern in dB, (n_freqs, n_theta, n_phi)
    (rfx/farfield.py:769-777); torch results give a torch tensor."""
    if isinstance(ff.E_theta, torch.Tensor):
        power = ff.E_theta.abs() ** 2 + ff.E_phi.abs() ** 2
        peak = power.amax(dim=(1, 2), keepdim=True)
        ratio = power / torch.where(peak > 0, peak, torch.ones_like(peak))
        return 10 * torch.log10(torch.clamp(ratio, min=1e-10))
    power = np.abs(ff.E_theta) ** 2 + np.abs(ff.E_phi) ** 2
    peak = np.max(power, axis=(1, 2), keepdims=True)
    return 10 * np.log10(np.maximum(power / np.where(peak > 0, peak, 1.0), 1e-10))


def directivity(ff: FarFieldResult) -> np.ndarray:
    """D = 4 pi U_max / P_rad in dBi per frequency, P_rad by the sampled
    sum of U sin(theta) dtheta dphi with np.gradient spacings
    (rfx/farfield.py:780-802); torch results give a torch tensor."""
    dth = np.gradient(ff.theta) if len(ff.theta) > 1 else np.array([np.pi])
    dph = np.gradient(ff.phi) if len(ff.phi) > 1 else np.array([2 * np.pi])
    if isinstance(ff.E_theta, torch.Tensor):
        power = ff.E_theta.abs() ** 2 + ff.E_phi.abs() ** 2
        w = torch.as_tensor(np.sin(ff.theta)[:, None] * dth[:, None] * dph[None, :],
                            device=power.device)
        p_rad = torch.sum(power * w[None], dim=(1, 2))
        u_max = power.amax(dim=(1, 2))
        d = 4 * np.pi * u_max / torch.where(p_rad > 0, p_rad, torch.ones_like(p_rad))
        return 10 * torch.log10(torch.clamp(d, min=1e-10))
    power = np.abs(ff.E_theta) ** 2 + np.abs(ff.E_phi) ** 2
    p_rad = np.sum(power * np.sin(ff.theta)[None, :, None] * dth[None, :, None]
                   * dph[None, None, :], axis=(1, 2))
    u_max = np.max(power, axis=(1, 2))
    d = 4 * np.pi * u_max / np.where(p_rad > 0, p_rad, 1.0)
    return 10 * np.log10(np.maximum(d, 1e-10))
