"""Per-cell material ids assembled on the GPU.

The host loops of the reference's material assembly run as HIP kernels over
the uint16 id array (include/rfxhip.h ``rfxh_assemble_*``):

* box fill in insertion order, later boxes winning
  (rfx/api/_compile.py:206-224; the node set of a box is the product of
  its three float32 axis masks, rfx/geometry/csg.py:235-307, each a
  contiguous index range);
* the CPML pad extension (rfx/geometry/rasterize_grid.py:237-414);
* the tangential PEC zero bits (rfx/boundaries/pec.py:90-176).

The material table stays on the host and gains its entries in the order the
host assembly (``MaterialMap``) creates them, so both give the same id array
bit for bit.  Between the fill and the PEC bits the caller may read and edit
single cells (ports: ``cell_props``, ``set_cell``, ``pec``) as on a
``MaterialMap``; ``finish`` then folds the PEC bits in and copies the ids to
the host map.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .materials import F32, MaterialMap


def mask_range(mask: np.ndarray):
    """[lo, hi) of a contiguous boolean axis mask (lo == hi when empty)."""
    nz = np.nonzero(mask)[0]
    if len(nz) == 0:
        return 0, 0
    lo, hi = int(nz[0]), int(nz[-1]) + 1
    if hi - lo != len(nz):
        raise ValueError("box axis mask is not one index range")
    return lo, hi


class _CellView:
    """Element access to a device array for the few cells the host edits."""

    def __init__(self, t: torch.Tensor):
        self.t = t

    def __getitem__(self, idx):
        return bool(self.t[tuple(idx)].item())

    def __setitem__(self, idx, value):
        self.t[tuple(idx)] = 1 if value else 0


class DeviceAssembly:
    """Material ids of one grid, built on ``device``; the table is ``mat``'s."""

    def __init__(self, mat: MaterialMap, device):
        _lib.require_device()
        self.mat = mat
        self.device = torch.device(device)
        self.lib = _lib.load()
        self.shape = tuple(int(n) for n in mat.shape)
        self.ids = torch.empty(self.shape, dtype=torch.int16, device=self.device)
        self.pec_t = torch.empty(self.shape, dtype=torch.uint8, device=self.device)
        self.any_pec = False
        self.pec = None          # set by fill(): a cell view when a PEC box exists

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _vacuum(self):
        eps, sig, mu, _ = self.mat.props()
        return np.ascontiguousarray(((eps == F32(1.0)) & (sig == F32(0.0))
                                     & (mu == F32(1.0))).astype(np.uint8))

    def fill(self, boxes, pads=None):
        """boxes: [(axis_masks, (eps_r, sigma, mu_r) or None for a PEC box)]
        in insertion order; pads: the six CPML pad widths, or None."""
        arr = (_lib.AsmBox * max(len(boxes), 1))()
        for q, (masks, props) in enumerate(boxes):
            rng = [mask_range(m) for m in masks]
            b = arr[q]
            for a in range(3):
                b.lo[a], b.hi[a] = rng[a]
            if props is None:
                b.pec, b.id = 1, 0
                self.any_pec = self.any_pec or all(h > l for l, h in rng)
            else:
                b.pec = 0
                # an empty box creates no table entry (MaterialMap.fill_box)
                b.id = self.mat.get_id(*props, 0) if all(h > l for l, h in rng) else 0
        vac = self._vacuum()
        pad_arr = None if pads is None else (C.c_int32 * 6)(*[int(p) for p in pads])
        nx, ny, nz = self.shape
        with torch.cuda.device(self.device):
            _lib.check(self.lib.rfxh_assemble_fill(
                nx, ny, nz, arr, len(boxes), pad_arr,
                vac.ctypes.data_as(C.POINTER(C.c_uint8)), len(vac),
                self.ids.data_ptr(), self.pec_t.data_ptr(), self._stream()))
        self.pec = _CellView(self.pec_t) if self.any_pec else None

    # -- single-cell access, as MaterialMap ------------------------------------
    def cell_props(self, idx):
        return self.mat.table[int(self.ids[tuple(idx)].item()) & 0xFFFF]

    def set_cell(self, idx, eps=None, sig=None, mu=None, add_flags=0):
        cur = int(self.ids[tuple(idx)].item()) & 0xFFFF
        mid = self.mat._derive(cur, eps, sig, mu, add_flags)
        self.ids[tuple(idx)] = int(np.uint16(mid).view(np.int16))

    # -- PEC bits and hand-over ----------------------------------------------------
    def finish(self) -> MaterialMap:
        """Fold the tangential PEC bits into the ids (new table entries in
        the order MaterialMap.add_flags creates them: by id, then bits) and
        copy the ids into the host map."""
        nx, ny, nz = self.shape
        if self.any_pec:
            n_mat = len(self.mat.table)
            used = torch.zeros(n_mat, dtype=torch.int32, device=self.device)
            with torch.cuda.device(self.device):
                _lib.check(self.lib.rfxh_assemble_pec_pairs(
                    nx, ny, nz, self.ids.data_ptr(), self.pec_t.data_ptr(), used.data_ptr(),
                    n_mat, self._stream()))
            used = used.cpu().numpy().view(np.uint32)
            table = np.zeros(n_mat * 8, np.uint16)
            for mid in np.nonzero(used)[0]:
                e, s, m, f = self.mat.table[int(mid)]
                for bits in range(1, 8):
                    if (int(used[mid]) >> bits) & 1:
                        table[int(mid) * 8 + bits] = self.mat.get_id(e, s, m, f | bits)
            with torch.cuda.device(self.device):
                _lib.check(self.lib.rfxh_assemble_pec_apply(
                    nx, ny, nz, self.ids.data_ptr(), self.pec_t.data_ptr(),
                    table.ctypes.data_as(C.POINTER(C.c_uint16)), n_mat, self._stream()))
        self.mat.ids = self.ids.cpu().numpy().view(np.uint16).reshape(self.shape)
        return self.mat
