"""Subpixel smoothing: per-component permittivity at interface cells.

``Simulation.run(subpixel_smoothing=True)`` (rfx/runners/uniform.py:271-278,
graded rfx/runners/nonuniform.py:648-672): each E component samples the
geometry at its own Yee position; at cells its sample point sees part-filled,
eps is the Kottke average of the box's eps and what the cell held before
(rfx/geometry/smoothing.py:244-293, 454-611, graded :296-452):

    f       = clip(1/2 - sdf / l, 0, 1)            (fill fraction, l the cell length)
    1/eps_c = n_c^2 (f/eps_in + (1-f)/eps_out) + (1 - n_c^2) / (f eps_in + (1-f) eps_out)

with the box's signed distance and its nearest-face normal n.  Boxes sharing
an eps_r are one group (union SDF, the normal of the nearest surface); groups
apply in order of first appearance.  Everything is float32, as the reference
computes it.  Only the cells within one cell of a group's boxes can change,
so each group is evaluated on that sub-block of the grid (the others keep
f = 0 and their value).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
_TINY = F32(1e-30)


def _box_terms(X, Y, Z, lo, hi):
    """Per-axis signed offsets from a box's faces, the box's signed
    distance and its nearest-face unit normal (float32)."""
    lo = np.asarray(lo, F32)
    hi = np.asarray(hi, F32)
    mid = (lo + hi) / F32(2.0)
    half = (hi - lo) / F32(2.0)
    r = [np.abs(X - mid[0]) - half[0], np.abs(Y - mid[1]) - half[1],
         np.abs(Z - mid[2]) - half[2]]
    out = np.sqrt(sum(np.maximum(v, F32(0.0)) ** 2 for v in r))
    ins = np.minimum(np.maximum(np.maximum(r[0], r[1]), r[2]), F32(0.0))
    sdf = out + ins
    along_x = (r[0] >= r[1]) & (r[0] >= r[2])
    along_y = (r[1] > r[0]) & (r[1] >= r[2])
    n = [np.where(along_x, np.sign(X - mid[0]), F32(0.0)),
         np.where(along_y, np.sign(Y - mid[1]), F32(0.0)),
         np.where(along_x | along_y, F32(0.0), np.sign(Z - mid[2]))]
    norm = np.sqrt(n[0] ** 2 + n[1] ** 2 + n[2] ** 2 + _TINY)
    return sdf.astype(F32), [(v / norm).astype(F32) for v in n]


def _index_window(coord, lo, hi, reach):
    """Indices whose coordinate lies within ``reach`` of [lo, hi]."""
    a = int(np.searchsorted(coord, F32(lo) - reach, side="left"))
    b = int(np.searchsorted(coord, F32(hi) + reach, side="right"))
    return max(a - 1, 0), min(b + 1, len(coord))


def kottke_eps(samples, length, shape, boxes, background=1.0):
    """Smoothed (eps_x, eps_y, eps_z).

    samples[c] = three 1-D float32 coordinate arrays (along x, y, z) of the
    E_c sample points; length: float32 scalar or a [nx][ny][nz]-broadcastable
    array (the cell length dividing the SDF); boxes = [(corner_lo,
    corner_hi, eps_r)] in insertion order."""
    out = [np.full(shape, F32(background), F32) for _ in range(3)]
    order, members = [], {}
    for lo, hi, e in boxes:
        key = float(e)
        if key not in members:
            order.append(key)
            members[key] = []
        members[key].append((lo, hi))
    length = np.asarray(length, F32)
    reach = F32(2.0) * F32(np.max(length))
    for e_in in order:
        group = members[e_in]
        for c in range(3):
            ax = samples[c]
            # sub-block: every cell within reach of a box of the group
            win = []
            for a in range(3):
                lo_i = min(_index_window(ax[a], b[0][a], b[1][a], reach)[0] for b in group)
                hi_i = max(_index_window(ax[a], b[0][a], b[1][a], reach)[1] for b in group)
                win.append(slice(lo_i, hi_i))
            if any(s.stop <= s.start for s in win):
                continue
            X = ax[0][win[0]][:, None, None]
            Y = ax[1][win[1]][None, :, None]
            Z = ax[2][win[2]][None, None, :]
            sub = tuple(s.stop - s.start for s in win)
            sdf = near = nrm = None
            for lo, hi in group:
                s_b, n_b = _box_terms(X, Y, Z, lo, hi)
                s_b = np.broadcast_to(s_b, sub)
                n_b = [np.broadcast_to(v, sub) for v in n_b]
                if sdf is None:
                    sdf, near, nrm = s_b, np.abs(s_b), n_b
                    continue
                sdf = np.minimum(sdf, s_b)
                closer = np.abs(s_b) < near
                near = np.where(closer, np.abs(s_b), near)
                nrm = [np.where(closer, u, v) for u, v in zip(n_b, nrm)]
            ln = length if length.ndim == 0 else np.broadcast_to(length, shape)[tuple(win)]
            f = np.clip(F32(0.5) - sdf / ln, F32(0.0), F32(1.0)).astype(F32)
            full = f >= F32(1.0)
            part = (f > F32(0.0)) & (f < F32(1.0))
            prev = out[c][tuple(win)]
            ei = F32(e_in)
            inv_n = f / ei + (F32(1.0) - f) / (prev + _TINY)
            inv_t = F32(1.0) / (f * ei + (F32(1.0) - f) * prev + _TINY)
            w = nrm[c] * nrm[c]
            avg = F32(1.0) / (w * inv_n + (F32(1.0) - w) * inv_t + _TINY)
            out[c][tuple(win)] = np.where(full, ei, np.where(part, avg, prev)).astype(F32)
    return tuple(out)


def smoothed_eps(grid, boxes):
    """Per-component eps_r of ``grid`` (uniform Grid or GradedGrid) for
    the boxes [(corner_lo, corner_hi, eps_r)]."""
    nodes = [np.asarray(v, F32) for v in grid.node_coords()]
    if getattr(grid, "d32", None) is None:
        # uniform: E_c half a cell along its own axis, SDF over dx
        half = F32(grid.dx * 0.5)
        centres = [v + half for v in nodes]
        length = F32(grid.dx)
    else:
        d = [np.asarray(v, F32) for v in grid.d32]
        centres = [nodes[a] + d[a] / F32(2.0) for a in range(3)]
        length = ((d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :])
                  ** F32(1.0 / 3.0)).astype(F32)
    samples = [[centres[a] if a == c else nodes[a] for a in range(3)] for c in range(3)]
    return kottke_eps(samples, length, tuple(grid.shape), boxes)
