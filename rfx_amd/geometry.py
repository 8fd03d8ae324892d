"""Axis-aligned Box and its rasterisation rule (rfx/geometry/csg.py:72-311).

A node belongs to the box iff lo <= x < hi per axis, compared in float32
against float32 node coordinates; an extent within 1.01 local cells takes
the thin-sheet branch and selects the single node nearest the centre
(csg.py:243-302).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def axis_mask(coords: np.ndarray, lo: float, hi: float) -> np.ndarray:
    coords = np.asarray(coords, dtype=np.float32)
    mid = (lo + hi) * 0.5
    extent = float(hi - lo)
    n = coords.size
    if n <= 1:
        dc = np.float32(1e-3)
    else:
        km = min(max(int(np.searchsorted(coords, np.float32(mid))) - 1, 0), n - 2)
        im1, ip1 = max(km - 1, 0), min(km + 1, n - 1)
        sl = coords[km] - coords[im1]
        sr = coords[ip1] - coords[km]
        dc = min(sl, sr) if (sl > 0 and sr > 0) else max(sl, sr)
    if np.float32(extent) <= np.float32(dc) * np.float32(1.01):
        m = np.zeros(n, bool)
        m[int(np.argmin(np.abs(coords - np.float32(mid))))] = True
        return m
    return (coords >= np.float32(lo)) & (coords < np.float32(hi))


@dataclass(frozen=True)
class Box:
    corner_lo: tuple
    corner_hi: tuple

    def bounding_box(self):
        return (self.corner_lo, self.corner_hi)

    def axis_masks(self, coords):
        return [axis_mask(coords[a], self.corner_lo[a], self.corner_hi[a]) for a in range(3)]

    def mask(self, grid) -> np.ndarray:
        mx, my, mz = self.axis_masks(grid.node_coords())
        return mx[:, None, None] & my[None, :, None] & mz[None, None, :]
