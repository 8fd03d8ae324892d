"""Per-cell material ids and the coefficient table the kernels read.

The reference keeps three float32 arrays (eps_r, sigma, mu_r) per cell and
recomputes ca/cb every step (rfx/core/yee.py:386-393).  Here every cell
holds a uint16 id into a table of distinct (eps_r, sigma, mu_r, flags)
entries; the table carries the precomputed float32 coefficients
  ca = (1 - L)/(1 + L), cb = (dt/eps)/(1 + L), L = sigma dt / (2 eps)
  ce = dt/(eps_r eps0)   (CPML E correction, rfx/boundaries/cpml.py:538-546)
  ch = dt/(mu_r mu0)     (H update and CPML H correction, yee.py:272, cpml.py:817)
computed with the reference's float32 operation order.  Flags mark cells
whose tangential E is zeroed by the PEC mask (rfx/boundaries/pec.py:90-208)
and cells that carry point operations (sources, probes, ports).
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from .constants import EPS_0, MU_0

F32 = np.float32
MAX_IDS = 65536
ANISO_FLAG = 16      # the id carries per-component eps (subpixel smoothing)
POLE_SHIFT = 16      # flag bit 16 + q: Debye pole q acts at the cell
MAX_POLES = 16


class DebyePole(NamedTuple):
    """One Debye relaxation term d_eps / (1 + j w tau) (rfx/materials/debye.py:38-48)."""
    delta_eps: float
    tau: float


class LorentzPole(NamedTuple):
    """One Lorentz oscillator kappa / (w0^2 - w^2 + 2 j delta w); Drude with
    omega_0 = 0 (rfx/materials/lorentz.py:31-43)."""
    omega_0: float
    delta: float
    kappa: float


def drude_pole(omega_p: float, gamma: float) -> LorentzPole:
    """Drude term from the plasma frequency and collision rate (rad/s)
    (rfx/materials/lorentz.py:46-56)."""
    return LorentzPole(omega_0=0.0, delta=gamma / 2.0, kappa=omega_p ** 2)


def lorentz_pole(delta_eps: float, omega_0: float, delta: float) -> LorentzPole:
    """Lorentz term from d_eps, resonance and damping (rad/s)
    (rfx/materials/lorentz.py:59-71)."""
    return LorentzPole(omega_0=omega_0, delta=delta, kappa=delta_eps * omega_0 ** 2)


class MaterialMap:
    def __init__(self, shape):
        self.shape = tuple(shape)
        self.ids = np.zeros(self.shape, np.uint16)
        self.table = []           # (eps_r, sigma, mu_r, flags)
        self.aniso = {}           # id -> (eps_x, eps_y, eps_z) of ANISO_FLAG ids
        self._lookup = {}
        self.get_id(1.0, 0.0, 1.0, 0)

    # -- table -------------------------------------------------------------
    def get_id(self, eps, sig, mu, flags=0, aniso=None) -> int:
        """Id of (eps_r, sigma, mu_r, flags[, per-component eps]); a new
        entry when the combination is new."""
        flags = int(flags)
        tail = b""
        if aniso is not None:
            flags |= ANISO_FLAG
            aniso = tuple(F32(v) for v in aniso)
            tail = b"".join(v.tobytes() for v in aniso)
        key = (F32(eps).tobytes(), F32(sig).tobytes(), F32(mu).tobytes(), flags, tail)
        mid = self._lookup.get(key)
        if mid is None:
            mid = len(self.table)
            if mid >= MAX_IDS:
                raise ValueError(
                    f"more than {MAX_IDS} distinct (eps_r, sigma, mu_r) combinations; "
                    "the material-id layout cannot represent this problem")
            self.table.append((F32(eps), F32(sig), F32(mu), flags))
            if aniso is not None:
                self.aniso[mid] = aniso
            self._lookup[key] = mid
        return mid

    def _derive(self, mid, eps=None, sig=None, mu=None, add_flags=0) -> int:
        """Id of material ``mid`` with some properties replaced and flags
        added (its per-component eps kept)."""
        e, s, m, f = self.table[int(mid)]
        return self.get_id(e if eps is None else eps, s if sig is None else sig,
                           m if mu is None else mu, f | int(add_flags),
                           self.aniso.get(int(mid)))

    def props(self):
        t = self.table
        eps = np.array([e[0] for e in t], F32)
        sig = np.array([e[1] for e in t], F32)
        mu = np.array([e[2] for e in t], F32)
        flg = np.array([e[3] for e in t], np.uint32)
        return eps, sig, mu, flg

    # -- geometry ------------------------------------------------------------
    def fill_box(self, axis_masks, eps, sig, mu):
        """Insertion-order fill: later shapes win (rfx/api/_compile.py:206-224)."""
        idx = [np.nonzero(m)[0] for m in axis_masks]
        if any(len(a) == 0 for a in idx):
            return
        mid = self.get_id(eps, sig, mu, 0)
        self.ids[np.ix_(*idx)] = mid

    def extend_pads(self, pads):
        """rfx/geometry/rasterize_grid.py:237-414 on ids.

        pads = (lo_x, hi_x, lo_y, hi_y, lo_z, hi_z).  Lo faces replicate the
        first interior slice outward; hi faces replicate the last interior
        slice, except where that slice is vacuum and the one inside it is
        not: there the inner slice is used and also written onto the last
        interior slice (the node the half-open box rule dropped).
        """
        eps, sig, mu, _ = self.props()
        for ax in range(3):
            plo, phi = pads[2 * ax], pads[2 * ax + 1]
            n = self.shape[ax]
            ids = np.moveaxis(self.ids, ax, 0)  # view
            if plo > 0:
                ids[:plo] = ids[plo:plo + 1]
            if phi > 0:
                outer = ids[n - phi - 1].copy()
                src = outer
                if n - plo - phi >= 2:
                    inner = ids[n - phi - 2]
                    vac_o = (eps[outer] == 1.0) & (sig[outer] == 0.0) & (mu[outer] == 1.0)
                    vac_i = (eps[inner] == 1.0) & (sig[inner] == 0.0) & (mu[inner] == 1.0)
                    src = np.where(vac_o & ~vac_i, inner, outer)
                    ids[n - phi - 1] = src
                ids[n - phi:] = src

    # -- material sweeps (rfx/vmap_sweep.py:265-390) ----------------------------
    _FIELD = {"eps_r": 0, "sigma": 1, "mu_r": 2}

    def set_field_where(self, mask, field: str, value):
        """Give the cells of ``mask`` the value of one property (eps_r, sigma
        or mu_r), keeping their other properties and flags."""
        where = np.nonzero(mask)
        if len(where[0]) == 0:
            return
        f = self._FIELD[field]
        uniq, inv = np.unique(self.ids[where], return_inverse=True)
        new = np.empty(len(uniq), np.uint16)
        for q, m in enumerate(uniq):
            props = list(self.table[int(m)])
            props[f] = F32(value)
            new[q] = self.get_id(*props)
        self.ids[where] = new[inv]

    def set_field_all(self, field: str, value):
        """Every material whose property differs from vacuum (eps_r, mu_r != 1,
        sigma != 0) takes ``value`` for it."""
        f = self._FIELD[field]
        vac = F32(0.0) if field == "sigma" else F32(1.0)
        table = []
        for e in self.table:
            e = list(e)
            if e[f] != vac:
                e[f] = F32(value)
            table.append(tuple(e))
        self.table = table
        self._lookup = {}
        for mid, (e, sg, mu, fl) in enumerate(table):
            tail = b"".join(v.tobytes() for v in self.aniso[mid]) if mid in self.aniso else b""
            key = (F32(e).tobytes(), F32(sg).tobytes(), F32(mu).tobytes(), int(fl), tail)
            self._lookup.setdefault(key, mid)

    # -- per-cell edits ------------------------------------------------------
    def cell_props(self, idx):
        e, s, m, f = self.table[int(self.ids[idx])]
        return e, s, m, f

    def set_cell(self, idx, eps=None, sig=None, mu=None, add_flags=0):
        self.ids[idx] = self._derive(self.ids[idx], eps, sig, mu, add_flags)

    def add_flags(self, cell_mask_bits: np.ndarray):
        """OR per-cell flag bits (integer array of the grid shape, up to 32
        bits) into the ids."""
        where = np.nonzero(cell_mask_bits)
        if len(where[0]) == 0:
            return
        cur = self.ids[where].astype(np.int64)
        bits = cell_mask_bits[where].astype(np.int64)
        key = cur * (1 << 32) + bits
        uniq, inv = np.unique(key, return_inverse=True)
        new = np.empty(len(uniq), np.uint16)
        for q, kv in enumerate(uniq):
            new[q] = self._derive(int(kv >> 32), add_flags=int(kv & 0xFFFFFFFF))
        self.ids[where] = new[inv]

    def apply_aniso(self, eps_x, eps_y, eps_z) -> int:
        """Give every cell whose per-component eps differs from its eps_r an
        id carrying the three values (subpixel smoothing); sigma, mu_r and
        flags stay the cell's.  Returns the number of such cells."""
        eps, _, _, _ = self.props()
        e_cell = eps[self.ids]
        where = np.nonzero((eps_x != e_cell) | (eps_y != e_cell) | (eps_z != e_cell))
        if len(where[0]) == 0:
            return 0
        rows = np.stack([self.ids[where].astype(np.uint32),
                         np.asarray(eps_x, F32)[where].view(np.uint32),
                         np.asarray(eps_y, F32)[where].view(np.uint32),
                         np.asarray(eps_z, F32)[where].view(np.uint32)], axis=1)
        uniq, inv = np.unique(rows, axis=0, return_inverse=True)
        new = np.empty(len(uniq), np.uint16)
        for q, (mid, ex, ey, ez) in enumerate(uniq):
            e, s, m, f = self.table[int(mid)]
            comp = tuple(np.uint32(v).view(F32) for v in (ex, ey, ez))
            new[q] = self.get_id(e, s, m, f, comp)
        self.ids[where] = new[np.asarray(inv).ravel()]
        return len(where[0])

    def override(self, eps_r=None, sigma=None, mu_r=None):
        """Replace per-cell material values wholesale (forward(eps_override=...),
        rfx/api/_execute.py:2800-2805), keeping each cell's flags."""
        eps, sig, mu, flg = self.props()
        cur = self.ids
        E = eps[cur] if eps_r is None else np.asarray(eps_r, F32).reshape(self.shape)
        S = sig[cur] if sigma is None else np.asarray(sigma, F32).reshape(self.shape)
        M = mu[cur] if mu_r is None else np.asarray(mu_r, F32).reshape(self.shape)
        Fl = flg[cur]
        # distinct (eps, sigma, mu, flags) tuples: one 1-D unique per
        # property (bit patterns), then one over the combined int64 key
        parts = [np.unique(a.view(np.uint32).ravel() if a.dtype == F32 else a.ravel(),
                           return_inverse=True) for a in (E, S, M, Fl)]
        if max(len(v) for v, _ in parts[:3]) > MAX_IDS or len(parts[3][0]) > 256:
            raise ValueError(f"override produces more than {MAX_IDS} distinct materials")
        key = np.zeros(E.size, np.int64)       # < 65536^3 * 256 < 2^63
        for vals, inv in parts:
            key = key * len(vals) + inv.ravel()
        ukey, inv = np.unique(key, return_inverse=True)
        if len(ukey) > MAX_IDS:
            raise ValueError(
                f"override produces {len(ukey)} distinct materials (max {MAX_IDS})")
        if self.aniso:
            raise NotImplementedError("overrides of subpixel-smoothed materials")
        self.table, self._lookup = [], {}
        self.get_id(1.0, 0.0, 1.0, 0)
        new = np.empty(len(ukey), np.uint16)
        for q, kv in enumerate(ukey):
            idx = []
            for vals, _ in reversed(parts):
                kv, r = divmod(int(kv), len(vals))
                idx.append(vals[r])
            fl, m, s, e = idx
            new[q] = self.get_id(np.uint32(e).view(F32), np.uint32(s).view(F32),
                                 np.uint32(m).view(F32), int(fl))
        self.ids = new[inv.ravel()].reshape(self.shape)

    def cell_arrays(self):
        """Dense float32 eps_r, sigma, mu_r (for small grids and tests)."""
        eps, sig, mu, _ = self.props()
        return eps[self.ids], sig[self.ids], mu[self.ids]

    # -- kernel table ----------------------------------------------------------
    @staticmethod
    def _ca_cb(eps_r, sig, dt32):
        """ca, cb of the E update in the reference's float32 order
        (rfx/core/yee.py:386-393, per component :880-895)."""
        eps = eps_r * F32(EPS_0)
        loss = sig * dt32 / (F32(2.0) * eps)
        return (F32(1.0) - loss) / (F32(1.0) + loss), (dt32 / eps) / (F32(1.0) + loss)

    def coefficient_table(self, dt: float):
        """Per id (ca, cb, ce, ch) and flags; ids with per-component eps
        carry the Ex coefficients here (the CPML ce keeps eps_r)."""
        eps_r, sig, mu_r, flg = self.props()
        dt32 = F32(dt)
        eps_x = eps_r.copy()
        for mid, comp in self.aniso.items():
            eps_x[mid] = comp[0]
        ca, cb = self._ca_cb(eps_x, sig, dt32)
        ce = dt32 / (eps_r * F32(EPS_0))
        ch = dt32 / (mu_r * F32(MU_0))
        coef = np.stack([ca, cb, ce, ch], axis=1).astype(F32)
        return np.ascontiguousarray(coef), np.ascontiguousarray(flg)

    def aniso_table(self, dt: float):
        """Per id (ca_y, cb_y, ca_z, cb_z) of the per-component ids, zero
        rows elsewhere; None when no id has per-component eps."""
        if not self.aniso:
            return None
        _, sig, _, _ = self.props()
        dt32 = F32(dt)
        rows = np.zeros((len(self.table), 4), F32)
        for mid, (_, ey, ez) in self.aniso.items():
            s = sig[mid:mid + 1]
            rows[mid, 0:2] = [v[0] for v in self._ca_cb(np.array([ey], F32), s, dt32)]
            rows[mid, 2:4] = [v[0] for v in self._ca_cb(np.array([ez], F32), s, dt32)]
        return np.ascontiguousarray(rows)


def dispersion_tables(mat: MaterialMap, dt: float, debye_poles, lorentz_poles=()):
    """Coefficient table and per-pole rows when dispersion poles exist, in
    the reference's float32 operation order.  Pole q (Debye first, then
    Lorentz) acts where the material's flags carry bit 16 + q.

    Debye (rfx/materials/debye.py:76-179): alpha = (2 tau - dt)/(2 tau + dt),
    beta = eps0 d_eps dt/(2 tau + dt); gamma = eps_inf + sum beta +
    sigma dt/2, ca = (eps_inf - sum beta - sigma dt/2)/gamma, cb = dt/gamma,
    cc = (1 - alpha)/gamma.  Lorentz (rfx/materials/lorentz.py:103-198):
    a = (2 - w0^2 dt^2)/(1 + delta dt), b = -(1 - delta dt)/(1 + delta dt),
    c = eps0 kappa dt^2/(1 + delta dt); gamma = eps_inf + sigma dt/2,
    ca = (eps_inf - sigma dt/2)/gamma, cb = dt/gamma, cc = 1/gamma.  Both
    (rfx/simulation.py:470-478): gamma_t = 1/cc_L + sum beta, ca = (ca_L/cc_L -
    sum beta)/gamma_t, cb = dt/gamma_t, cc_D = (1 - alpha)/gamma_t,
    cc_L = 1/gamma_t.  Gammas are floored at eps0 1e-10.

    Returns (coef [n_mat][4] with the dispersive ca, cb and the usual ce,
    ch; flags; rows [n_mat][n_pole][4] = Debye (alpha, beta, cc, 0),
    Lorentz (a, b, c, cc))."""
    coef, flg = mat.coefficient_table(dt)
    eps_r, sig, _, _ = mat.props()
    n_mat, nd, nl = len(flg), len(debye_poles), len(lorentz_poles)
    dt32 = F32(dt)
    floor = F32(EPS_0 * 1e-10)

    def on(q):
        return ((flg >> (POLE_SHIFT + q)) & 1).astype(bool)

    eps_inf = eps_r * F32(EPS_0)
    half_loss = sig * dt32 / F32(2.0)
    rows = np.zeros((n_mat, nd + nl, 4), F32)
    alpha, beta = [], []
    for q, (de, tau) in enumerate(debye_poles):
        alpha.append(np.where(on(q), F32((2.0 * tau - dt) / (2.0 * tau + dt)), F32(0.0)))
        beta.append(np.where(on(q), F32(EPS_0 * de * dt / (2.0 * tau + dt)), F32(0.0)))
    beta_sum = np.zeros(n_mat, F32)
    for b in beta:
        beta_sum = beta_sum + b
    for l, (w0, delta, kappa) in enumerate(lorentz_poles):
        denom = 1.0 + delta * dt
        m = on(nd + l)
        rows[:, nd + l, 0] = np.where(m, F32((2.0 - w0 ** 2 * dt ** 2) / denom), F32(0.0))
        rows[:, nd + l, 1] = np.where(m, F32(-(1.0 - delta * dt) / denom), F32(0.0))
        rows[:, nd + l, 2] = np.where(m, F32(EPS_0 * kappa * dt ** 2 / denom), F32(0.0))
    coef = coef.copy()
    if nl == 0:
        gamma = np.maximum(eps_inf + beta_sum + half_loss, floor)
        coef[:, 0] = (eps_inf - beta_sum - half_loss) / gamma
        coef[:, 1] = dt32 / gamma
    else:
        g_l = np.maximum(eps_inf + half_loss, floor)
        ca_l = (eps_inf - half_loss) / g_l
        cc_l = F32(1.0) / g_l
        if nd == 0:
            gamma = g_l
            coef[:, 0] = ca_l
            coef[:, 1] = dt32 / g_l
        else:
            g_base = F32(1.0) / cc_l
            gamma = np.maximum(g_base + beta_sum, floor)
            coef[:, 0] = (ca_l * g_base - beta_sum) / gamma
            coef[:, 1] = dt32 / gamma
            cc_l = F32(1.0) / gamma
        for l in range(nl):
            rows[:, nd + l, 3] = cc_l
    for q in range(nd):
        rows[:, q, 0] = alpha[q]
        rows[:, q, 1] = beta[q]
        rows[:, q, 2] = (F32(1.0) - alpha[q]) / gamma
    return np.ascontiguousarray(coef), flg, np.ascontiguousarray(rows)


def tangential_edge_bits(pec_mask: np.ndarray) -> np.ndarray:
    """uint8 bits 1/2/4 = Ex/Ey/Ez zeroed (rfx/boundaries/pec.py:90-176):
    a component is zeroed at a PEC cell iff a neighbour along that
    component's axis is PEC too (zero padding past the domain)."""
    bits = np.zeros(pec_mask.shape, np.uint8)
    for ax in range(3):
        nb = np.zeros_like(pec_mask)
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(1, None)
        sl_b[ax] = slice(0, -1)
        nb[tuple(sl_a)] |= pec_mask[tuple(sl_b)]
        nb[tuple(sl_b)] |= pec_mask[tuple(sl_a)]
        bits |= ((pec_mask & nb).astype(np.uint8) << ax)
    return bits
