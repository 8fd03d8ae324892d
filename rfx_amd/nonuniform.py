"""Graded Yee grid: host set-up of the non-uniform lane.

Restates rfx/nonuniform.py:37-422 (NonUniformGrid, make_nonuniform_grid,
_profile_to_inv_arrays, position_to_index) and the source / port metrics
of rfx/nonuniform.py:588-691 and rfx/runners/nonuniform.py:694-887 for the
HIP kernels, which take per-index inverse spacings (``inv_e`` for the E
update, ``inv_h`` for the H update) instead of one dx.

Each axis is the physical profile padded by ``cpml_layers`` copies of its
boundary cell per face plus one duplicated bounding cell (N cells -> N+1
E nodes).  The reference keeps the padded profiles as float32 arrays and
derives every metric from them (inverse spacings, node coordinates, index
lookup, source volumes, port metrics); dt comes from the float64 profile
minimum with c = 1/sqrt(eps0 mu0).
"""
from __future__ import annotations

import numpy as np

from .constants import C0, EPS_0, MU_0

F32 = np.float32
C0_NU = 1.0 / np.sqrt(float(EPS_0) * float(MU_0))   # rfx/nonuniform.py:34
AXIS = {"ex": 0, "ey": 1, "ez": 2}


class GradedGrid:
    """Graded Yee grid: per-axis physical cell-size profiles; an axis given
    no profile is uniform with spacing dx over its domain extent
    (rfx/runners/nonuniform.py:25-60, rfx/nonuniform.py:274-422)."""

    def __init__(self, freq_max: float, domain, dz_profile=None, dx: float | None = None,
                 cpml_layers: int = 8, dx_profile=None, dy_profile=None):
        self.freq_max = float(freq_max)
        if (dx_profile is not None or dy_profile is not None) and dx is None:
            raise ValueError("dx_profile / dy_profile require an explicit dx (boundary cell size)")
        self.dx = float(dx) if dx is not None else C0 / self.freq_max / 20.0
        p = int(cpml_layers)
        self.cpml_layers = p

        def prof(given, axis):
            if given is None:
                return np.full(max(1, int(round(float(domain[axis]) / self.dx))), self.dx)
            a = np.asarray(given, np.float64)
            if a.ndim != 1 or a.size < 1 or np.any(a <= 0):
                raise ValueError(f"{'xyz'[axis]} profile must be a non-empty 1-D array of "
                                 "positive cell sizes")
            return a

        px, py, pz = prof(dx_profile, 0), prof(dy_profile, 1), prof(dz_profile, 2)
        # the CPML uses the boundary cell sizes (rfx/nonuniform.py:323-360)
        if dx_profile is not None and (abs(px[0] - self.dx) > 1e-12 or abs(px[-1] - self.dx) > 1e-12):
            raise ValueError(f"dx_profile must start and end with the boundary dx={self.dx}")
        self.dy = float(py[0]) if dy_profile is not None else self.dx
        if dy_profile is not None and abs(py[-1] - self.dy) > 1e-12:
            raise ValueError("dy_profile boundary cells must match each other")
        self.d64 = [self._pad(a, p) for a in (px, py, pz)]
        self.d32 = [d.astype(F32) for d in self.d64]
        self.shape = tuple(len(d) for d in self.d64)
        self.nx, self.ny, self.nz = self.shape
        inv_sq = sum(1.0 / float(np.min(d)) ** 2 for d in self.d64)
        self.dt = float(0.99 / (C0_NU * np.sqrt(inv_sq)))
        self.axis_pads = (p, p, p)
        self.face_pads = (p,) * 6
        self.cpml_axes = "xyz"
        self.is_2d = False

    @staticmethod
    def _pad(prof, p):
        body = np.concatenate([np.full(p, float(prof[0])), np.asarray(prof, np.float64),
                               np.full(p, float(prof[-1]))])
        return np.concatenate([body, body[-1:]])

    def num_timesteps(self, num_periods: float = 20.0) -> int:
        return int(np.ceil(num_periods * (1.0 / self.freq_max) / self.dt))

    # ---- metrics ------------------------------------------------------------
    def inverse_spacings(self):
        """(inv_e, inv_h) per axis (rfx/nonuniform.py:187-224): H update
        1/d[k] with 0 at the last index; E update 2/(d[k-1]+d[k]) with
        1/d[0] at the first index, in float32."""
        inv_e, inv_h = [], []
        for d in self.d32:
            local = F32(1.0) / d
            inv_h.append(np.concatenate([local[:-1], np.zeros(1, F32)]))
            inv_e.append(np.concatenate([local[:1], F32(2.0) / (d[:-1] + d[1:])]))
        return inv_e, inv_h

    @property
    def dz(self):
        """Padded float32 z cell sizes, bounding cell included
        (NonUniformGrid.dz, rfx/nonuniform.py:58)."""
        return self.d32[2]

    def face_cell_sizes(self):
        """CPML face cell sizes (rfx/boundaries/cpml.py:174-197): the scalar
        dx, dy and the float32 first / last dz entries."""
        dz = self.d32[2]
        return [(self.dx, self.dx), (self.dy, self.dy), (float(dz[0]), float(dz[-1]))]

    def node_coords(self):
        """float32 E-node coordinates, origin at the first interior node:
        float64 running sums of the float32 cell sizes
        (rfx/geometry/rasterize_grid.py:60-116)."""
        out = []
        for a, d in enumerate(self.d32):
            e = np.concatenate([[0.0], np.cumsum(d.astype(np.float64))])
            out.append((e[:-1] - e[self.axis_pads[a]]).astype(F32))
        return out

    def position_to_index(self, pos):
        """Nearest interior cell edge per axis (rfx/nonuniform.py:426-517):
        edges are the float32 running sum of the float32 interior cells and
        the distance is taken in float32; ties go to the lower index."""
        idx = []
        for a in range(3):
            d = self.d32[a]
            p = self.axis_pads[a]
            interior = d[p:len(d) - p - 1]
            edges = np.concatenate([np.zeros(1, F32), np.cumsum(interior, dtype=F32)])
            idx.append(int(np.argmin(np.abs(edges - F32(pos[a])))) + p)
        return tuple(idx)

    def dual(self, axis: int, k: int) -> float:
        """Dual spacing of E node k, float64 of the float32 cells
        (e_node_dual_spacing_at, rfx/nonuniform.py:253-271)."""
        d = self.d32[axis]
        return float(d[0]) if k <= 0 else 0.5 * (float(d[k - 1]) + float(d[k]))

    def dual_z_f32(self, k: int) -> float:
        """The z dual spacing as the port set-up evaluates it, in float32
        arithmetic on the float32 dz array (rfx/runners/nonuniform.py:773)."""
        d = self.d32[2]
        return float(d[0]) if k <= 0 else float(F32(0.5) * (d[k - 1] + d[k]))

    def primal(self, axis: int, k: int) -> float:
        return float(self.d32[axis][k])

    def port_metrics(self, cell, axis: int):
        """(d_parallel, dual_perp1, dual_perp2, (dual_x, dual_y, dual_z)) of a
        port cell (rfx/runners/nonuniform.py:766-796, 846-868)."""
        i, j, k = cell
        loop = (self.dual(0, i), self.dual(1, j), self.dual_z_f32(k))
        if axis == 2:
            return self.primal(2, k), loop[0], loop[1], loop
        if axis == 1:
            return self.primal(1, j), loop[0], loop[2], loop
        return self.primal(0, i), loop[1], loop[2], loop

    def __repr__(self):
        return (f"GradedGrid(shape={self.shape}, dx={self.dx:.4e} m, "
                f"dz=[{float(self.d64[2].min()):.3e}, {float(self.d64[2].max()):.3e}] m, "
                f"dt={self.dt:.4e} s)")


def make_nonuniform_grid(freq_max, domain, dz_profile=None, dx=None, cpml_layers=8,
                         dx_profile=None, dy_profile=None) -> GradedGrid:
    return GradedGrid(freq_max, domain, dz_profile, dx, cpml_layers, dx_profile, dy_profile)


def current_source_scale(grid: GradedGrid, eps_r: float, sigma: float, idx, component: str,
                         dt: float) -> np.float32:
    """f32(Cb / dV) of a current source (rfx/nonuniform.py:588-691): Cb from
    the cell's eps_r / sigma in float64, dV = primal width on the
    component's own axis x dual spacings on the two transverse axes."""
    if component not in AXIS:
        raise ValueError(f"unknown source component {component!r}")
    eps = float(eps_r) * EPS_0
    loss = float(sigma) * dt / (2.0 * eps)
    cb = (dt / eps) / (1.0 + loss)
    a = AXIS[component]
    w = [grid.primal(x, idx[x]) if x == a else grid.dual(x, idx[x]) for x in range(3)]
    return F32(cb / (w[0] * w[1] * w[2]))
