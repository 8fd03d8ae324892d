"""Uniform Yee grid (mirrors rfx/grid.py:11-211 for the FDTD lane).

Shape per axis = ceil(L/dx) + 1 + pad_lo + pad_hi (rfx/grid.py:150-164);
dt = dx / (C0 sqrt(3)) * 0.99 (:94-96); position -> index uses Python's
round (half-to-even) plus the lo pad (:190-211).
"""
from __future__ import annotations

import numpy as np

from .constants import C0


class Grid:
    def __init__(self, freq_max: float, domain, dx: float | None = None,
                 cpml_layers: int = 8, cpml_axes: str = "xyz"):
        self.freq_max = float(freq_max)
        self.domain = tuple(float(d) for d in domain)
        self.cpml_layers = int(cpml_layers)
        self.cpml_axes = "".join(a for a in "xyz" if a in cpml_axes)
        self.dx = float(dx) if dx is not None else C0 / self.freq_max / 20.0
        self.dt = self.dx / (C0 * np.sqrt(3.0)) * 0.99
        pads = [self.cpml_layers if a in self.cpml_axes else 0 for a in "xyz"]
        self.pad_x_lo = self.pad_x_hi = pads[0]
        self.pad_y_lo = self.pad_y_hi = pads[1]
        self.pad_z_lo = self.pad_z_hi = pads[2]
        self.axis_pads = tuple(pads)
        self.face_pads = (pads[0], pads[0], pads[1], pads[1], pads[2], pads[2])
        self.nx, self.ny, self.nz = (int(np.ceil(d / self.dx)) + 1 + 2 * p
                                     for d, p in zip(self.domain, pads))
        self.shape = (self.nx, self.ny, self.nz)
        self.is_2d = False

    @staticmethod
    def courant_dt(dx: float, ndim: int = 3) -> float:
        return dx / (C0 * np.sqrt(float(ndim))) * 0.99

    def num_timesteps(self, num_periods: float = 20.0) -> int:
        return int(np.ceil(num_periods * (1.0 / self.freq_max) / self.dt))

    def position_to_index(self, pos):
        idx = tuple(int(round(p / self.dx)) + self.axis_pads[a] for a, p in enumerate(pos))
        if not all(0 <= idx[a] < self.shape[a] for a in range(3)):
            raise ValueError(f"position {pos} maps to grid index {idx}, outside the grid "
                             f"shape {self.shape}. Check the position lies inside the "
                             "simulation domain.")
        return idx

    def node_coords(self):
        """float32 node coordinates, f32(f32(i - pad) * f32(dx)) as the
        reference's rasteriser sees them (rfx/geometry/csg.py:61-69)."""
        return [(np.arange(n) - p).astype(np.float32) * np.float32(self.dx)
                for n, p in zip(self.shape, self.axis_pads)]

    # spacing consumed by the kernels: uniform -> one inverse length
    def inverse_spacings(self):
        inv = np.float32(1.0) / np.float32(self.dx)
        e = [np.full(n, inv, np.float32) for n in self.shape]
        return e, [a.copy() for a in e]

    def face_cell_sizes(self):
        return [(self.dx, self.dx)] * 3

    def __repr__(self):
        return (f"Grid(shape={self.shape}, dx={self.dx:.4e} m, dt={self.dt:.4e} s, "
                f"freq_max={self.freq_max:.2e} Hz, cpml_axes={self.cpml_axes!r})")
