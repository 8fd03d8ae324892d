"""numpy restatement of rfx's FDTD time-stepping path — TEST INFRASTRUCTURE ONLY.

Used by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` as the checker; never imported by the product package.

Every function names the reference file:line it restates (paths relative to
the rfx repository root).  Stepping arithmetic runs in float64 by default
(``dtype=np.float64``), the "FP64 run" that float32 results are quoted
against; ``dtype=np.float32`` gives a float32 run for bit-pattern studies.

Set-up quantities keep the reference's own float32 rounding wherever the
reference computes them in float32 even under x64 (material arrays, CPML
b/c/kappa, source waveform tables, node coordinates): the reference's
float64 mode keeps those float32 too (rfx/api/__init__.py:320-324,
rfx/simulation.py:273 ``jnp.arange(n_steps, dtype=jnp.float32)``).

Pinning (what checks each function; see tests/test_oracle_*.py):
  * uniform step, CPML, lumped-port fold, V/I DFT, S11 extraction:
    golden vector tests/fixtures/golden_forward_no_rlc_s11.npy (rtol 1e-5).
  * soft J-source, eps override, probe series, forward lane (no outer PEC):
    tests/fixtures/gradient_dx_ladder/ad_fd_ladder.json J0 / J+ / J- values.
  * DFT plane accumulator: equality with a direct DFT of the stored probe
    series at the same cell (closed form).
  * Box rasterisation: brute-force per-node comparison on tiny grids and
    the README quick-start footprint.
  * CPML profile: closed forms of rfx/boundaries/cpml.py:101-171
    (sigma_max formula, b = exp(...), c = sigma (b-1)/(sigma k + k^2 a)).
  * graded mesh: a uniform profile must reproduce the uniform lane after the
    documented source-normalisation factor; PEC-cavity resonance closed
    form.
  * Debye / Lorentz ADE (debye_coefficients, lorentz_coefficients, the
    dispersive Stepper.update_e): the reference's hand-computed
    coefficients; tau -> 0 / tau -> inf limits (a plain dielectric of
    eps_inf + d_eps / eps_inf); the exact polariton frequencies of a
    Lorentz-filled PEC cavity; kappa = 0 reducing the mixed update to the
    Debye one.
  * NTFF box (_ntff, far_field): faces equal time-shifted DFT planes;
    single-cell current elements (hand-derived n x H, -n x E); the
    short-dipole directivity the reference's own test records (1.84 dBi,
    4.78 dBi); far-field power against closed flux-monitor boxes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# rfx/grid.py:8
C0 = 299_792_458.0
# rfx/core/yee.py:117-118
EPS_0 = 8.8541878128e-12
MU_0 = 1.25663706212e-6
# rfx/nonuniform.py:34 (speed of light the graded mesh derives from eps0/mu0)
C0_NU = 1.0 / np.sqrt(float(EPS_0) * float(MU_0))

F32 = np.float32
COMPS = ("ex", "ey", "ez", "hx", "hy", "hz")


# ---------------------------------------------------------------------------
# Grid (rfx/grid.py:11-211)
# ---------------------------------------------------------------------------

class UniformGrid:
    """Uniform grid: shape incl. CPML pads, dt, position->index.

    rfx/grid.py:91-96 (dx default lambda/20, dt = dx/(C0 sqrt 3) * 0.99),
    :150-164 (n = ceil(L/dx) + 1 + pad_lo + pad_hi), :190-211 (index =
    Python round(pos/dx) + pad_lo, round-half-even).
    """

    def __init__(self, freq_max, domain, dx=None, cpml_layers=8):
        self.freq_max = float(freq_max)
        self.domain = tuple(float(d) for d in domain)
        self.dx = float(dx) if dx is not None else C0 / freq_max / 20.0
        self.dt = self.dx / (C0 * np.sqrt(3.0)) * 0.99
        self.pad = int(cpml_layers)
        self.cpml_layers = int(cpml_layers)
        self.shape = tuple(int(np.ceil(d / self.dx)) + 1 + 2 * self.pad
                           for d in self.domain)

    def position_to_index(self, pos):
        idx = tuple(int(round(p / self.dx)) + self.pad for p in pos)
        for a in range(3):
            if not 0 <= idx[a] < self.shape[a]:
                raise ValueError(f"position {pos} maps outside the grid")
        return idx

    def num_timesteps(self, num_periods=20.0):
        # rfx/grid.py:185-188
        return int(np.ceil(num_periods * (1.0 / self.freq_max) / self.dt))

    def node_coords_f32(self):
        """rfx/geometry/csg.py:61-69: f32(f32(i - pad) * f32(dx))."""
        return [(np.arange(n) - self.pad).astype(F32) * F32(self.dx)
                for n in self.shape]

    # per-axis spacing seen by the stepper
    def spacing(self):
        return dict(kind="uniform", dx=self.dx)


class GradedGrid:
    """Graded (non-uniform) grid: rfx/nonuniform.py:187-224 and 274-422.

    Profiles are the PHYSICAL interior cell sizes; each axis is padded by
    ``cpml_layers`` boundary-size cells per face (_pad_profile :98-128)
    and gets one duplicated bounding cell (_append_bounding_node :130-155).
    dt = 0.99 / (c sqrt(sum 1/dmin^2)) with c = 1/sqrt(eps0 mu0) (:386-401),
    from the float64 padded profiles.  Everything downstream of the grid
    reads the float32 cell-size arrays (``dx_arr``/``dy_arr``/``dz`` are
    jnp float32, :403-405): inverse spacings, node coordinates, index
    lookup, source volumes and port metrics.
    """

    def __init__(self, dz_profile, dx, domain_xy, cpml_layers=8,
                 dx_profile=None, dy_profile=None):
        p = int(cpml_layers)
        self.pad = p
        self.cpml_layers = p

        def _full(prof):
            prof = np.asarray(prof, dtype=np.float64)
            lo = np.full(p, float(prof[0]))
            hi = np.full(p, float(prof[-1]))
            full = np.concatenate([lo, prof, hi])
            return np.concatenate([full, full[-1:]])

        if dx_profile is None:
            dx_profile = np.full(int(round(domain_xy[0] / dx)), float(dx))
        if dy_profile is None:
            dy_profile = np.full(int(round(domain_xy[1] / dx)), float(dx))
        self.d_full = [_full(dx_profile), _full(dy_profile), _full(dz_profile)]
        self.d32 = [d.astype(F32) for d in self.d_full]
        self.shape = tuple(len(d) for d in self.d_full)
        dmins = [float(np.min(d)) for d in self.d_full]
        self.dt = float(0.99 / (C0_NU * np.sqrt(sum(1.0 / m ** 2 for m in dmins))))
        self.dx = float(dx)
        self.dy = float(np.asarray(dy_profile)[0])
        # rfx/nonuniform.py:212-224 (float32 arrays)
        self.inv_e, self.inv_h = [], []
        for a in self.d32:
            inv_local = F32(1.0) / a
            inv_mean = F32(2.0) / (a[:-1] + a[1:])
            self.inv_h.append(np.concatenate([inv_local[:-1], np.zeros(1, F32)]))
            self.inv_e.append(np.concatenate([inv_local[:1], inv_mean]))
        # CPML face cell sizes (rfx/boundaries/cpml.py:174-197): the scalar
        # boundary dx / dy and the float32 first/last dz entries
        dz = self.d32[2]
        self.face_d = {"x": (self.dx, self.dx), "y": (self.dy, self.dy),
                       "z": (float(dz[0]), float(dz[-1]))}

    def num_timesteps(self, freq_max, num_periods=20.0):
        return int(np.ceil(num_periods * (1.0 / freq_max) / self.dt))

    def node_positions(self, axis):
        """rfx/geometry/rasterize_grid.py:60-116: float64 cumulative sums of
        the float32 cell sizes, origin at the first interior node, cast to
        float32."""
        d = self.d32[axis].astype(np.float64)
        edges = np.concatenate([[0.0], np.cumsum(d)])
        return (edges[:-1] - edges[self.pad]).astype(F32)

    def node_coords_f32(self):
        return [self.node_positions(a) for a in range(3)]

    def position_to_index(self, pos):
        """rfx/nonuniform.py:426-517: nearest interior cell edge, with edges
        the float32 running sum of the float32 interior cells (numpy cumsum
        keeps float32) and the distance taken in float32; first minimum."""
        idx = []
        for a in range(3):
            d = self.d32[a]
            interior = d[self.pad:len(d) - self.pad - 1]
            edges = np.insert(np.cumsum(interior), 0, 0.0)
            dist = np.abs(edges - F32(pos[a]))
            idx.append(int(np.argmin(dist)) + self.pad)
        return tuple(idx)

    def dual64(self, axis, k):
        """e_node_dual_spacing_at on the float64 view of the float32 profile
        (rfx/nonuniform.py:253-271 with np.float64 input)."""
        d = self.d32[axis].astype(np.float64)
        return float(d[0]) if k <= 0 else float(0.5 * (d[k - 1] + d[k]))

    def dual_z32(self, k):
        """The same on the jnp float32 ``grid.dz`` (float32 arithmetic), as
        the port metrics read it (rfx/runners/nonuniform.py:773, 854)."""
        d = self.d32[2]
        return float(d[0]) if k <= 0 else float(F32(0.5) * (d[k - 1] + d[k]))

    def spacing(self):
        return dict(kind="graded", inv_e=self.inv_e, inv_h=self.inv_h)


# ---------------------------------------------------------------------------
# Geometry and materials (rfx/geometry/csg.py, rfx/api/_compile.py,
# rfx/geometry/rasterize_grid.py)
# ---------------------------------------------------------------------------

def _axis_mask(coords, lo, hi):
    """rfx/geometry/csg.py:243-302 (float32 comparisons, thin-sheet branch)."""
    coords = np.asarray(coords, dtype=F32)
    mid = (lo + hi) * 0.5
    extent = float(hi - lo)
    n = coords.size
    if n <= 1:
        dc_local = F32(1e-3)
    else:
        k_mid = int(np.clip(np.searchsorted(coords, F32(mid)) - 1, 0, n - 2))
        im1 = int(np.clip(k_mid - 1, 0, n - 1))
        ip1 = int(np.clip(k_mid + 1, 0, n - 1))
        s_left = coords[k_mid] - coords[im1]
        s_right = coords[ip1] - coords[k_mid]
        if s_left > 0 and s_right > 0:
            dc_local = min(s_left, s_right)
        else:
            dc_local = max(s_left, s_right)
    nearest = int(np.argmin(np.abs(coords - F32(mid))))
    thin = np.zeros(n, dtype=bool)
    thin[nearest] = True
    vol = (coords >= F32(lo)) & (coords < F32(hi))
    is_thin = F32(extent) <= F32(dc_local) * F32(1.01)
    return thin if is_thin else vol


def box_mask(coords, corner_lo, corner_hi):
    """rfx/geometry/csg.py:296-302: separable per-axis masks, outer-ANDed."""
    mx = _axis_mask(coords[0], corner_lo[0], corner_hi[0])
    my = _axis_mask(coords[1], corner_lo[1], corner_hi[1])
    mz = _axis_mask(coords[2], corner_lo[2], corner_hi[2])
    return mx[:, None, None] & my[None, :, None] & mz[None, None, :]


def extend_cpml_pad(eps_r, sigma, mu_r, p):
    """rfx/geometry/rasterize_grid.py:237-414 with equal pads on all faces.

    Lo faces replicate the first interior slice outward; hi faces replicate
    the last interior slice, falling back one slice inward (and repairing
    the dropped boundary node) where the last slice is vacuum but the one
    inside it is not.
    """
    arrays = [eps_r.copy(), sigma.copy(), mu_r.copy()]
    if p <= 0:
        return tuple(arrays)
    for ax in range(3):
        n = arrays[0].shape[ax]

        def sl(a, b):
            s = [slice(None)] * 3
            s[ax] = slice(a, b)
            return tuple(s)

        # lo face
        for i, a in enumerate(arrays):
            src = a[sl(p, p + 1)]
            a[sl(0, p)] = src
        # hi face
        outer, inner, dst = sl(n - p - 1, n - p), sl(n - p - 2, n - p - 1), sl(n - p, n)
        e, s, m = arrays
        outer_vac = (e[outer] == 1.0) & (s[outer] == 0.0) & (m[outer] == 1.0)
        use_inner = None
        if n - 2 * p >= 2:
            inner_vac = (e[inner] == 1.0) & (s[inner] == 0.0) & (m[inner] == 1.0)
            use_inner = outer_vac & ~inner_vac
        for a in arrays:
            src = a[outer].copy()
            if use_inner is not None:
                src = np.where(use_inner, a[inner], src)
                a[outer] = src
            a[dst] = src
    return tuple(arrays)


PEC_SIGMA_THRESHOLD = 1e6  # rfx/api/_compile.py:118


def assemble_materials(shape, coords, geometry, materials, cpml_pad, sweep=None):
    """rfx/api/_compile.py:121-367 (no dispersion / thin conductors).

    geometry: list of (corner_lo, corner_hi, material_name) in insertion
    order (later wins); materials: name -> (eps_r, sigma, mu_r).
    Returns float32 eps_r, sigma, mu_r and a bool pec_mask or None.

    sweep = (name or None, field, value): one element of a material sweep
    (rfx/vmap_sweep.py:265-390 _build_batched_materials).  With a material
    name, every cell inside that material's shapes (later shapes of other
    materials included) takes the value before the CPML pads are extended
    (:328-373); without one, every cell whose field differs from vacuum
    (eps_r, mu_r != 1, sigma != 0) takes it after the extension (:374-401).
    """
    eps_r = np.ones(shape, F32)
    sigma = np.zeros(shape, F32)
    mu_r = np.ones(shape, F32)
    pec = np.zeros(shape, bool)
    for lo, hi, name in geometry:
        e, s, m = materials[name]
        # box_mask is an outer product of per-axis masks, so the reference's
        # jnp.where(mask, value, array) equals assignment on the index box
        sel = np.ix_(*[np.nonzero(_axis_mask(coords[a], lo[a], hi[a]))[0]
                       for a in range(3)])
        if s >= PEC_SIGMA_THRESHOLD:
            pec[sel] = True
        else:
            eps_r[sel] = F32(e)
            sigma[sel] = F32(s)
            mu_r[sel] = F32(m)
    arrays = {"eps_r": eps_r, "sigma": sigma, "mu_r": mu_r}
    if sweep is not None and sweep[0] is not None:
        name, fld, value = sweep
        mask = np.zeros(shape, bool)
        for lo, hi, nm in geometry:
            if nm == name:
                mask[np.ix_(*[np.nonzero(_axis_mask(coords[a], lo[a], hi[a]))[0]
                              for a in range(3)])] = True
        arrays[fld] = np.where(mask, F32(value), arrays[fld]).astype(F32)
    eps_r, sigma, mu_r = arrays["eps_r"], arrays["sigma"], arrays["mu_r"]
    if cpml_pad > 0:
        eps_r, sigma, mu_r = extend_cpml_pad(eps_r, sigma, mu_r, cpml_pad)
    if sweep is not None and sweep[0] is None:
        _, fld, value = sweep
        arr = {"eps_r": eps_r, "sigma": sigma, "mu_r": mu_r}[fld]
        vac = F32(0.0) if fld == "sigma" else F32(1.0)
        arr = np.where(arr != vac, F32(value), arr).astype(F32)
        if fld == "eps_r":
            eps_r = arr
        elif fld == "sigma":
            sigma = arr
        else:
            mu_r = arr
    return eps_r, sigma, mu_r, (pec if pec.any() else None)


def lorentz_coefficients(poles, masks, eps_r, sigma, dt, dtype=np.float64):
    """rfx/materials/lorentz.py:103-198 (init_lorentz): per pole, inside its
    mask, a = (2 - w0^2 dt^2)/(1 + delta dt), b = -(1 - delta dt)/(1 +
    delta dt), c = eps0 kappa dt^2/(1 + delta dt); gamma = eps0 eps_inf +
    sigma dt/2, ca = (eps0 eps_inf - sigma dt/2)/gamma, cb = dt/gamma,
    cc = 1/gamma."""
    dt_ = dtype(dt)
    eps_inf = eps_r.astype(dtype) * dtype(EPS_0)
    sig = sigma.astype(dtype)
    a_l, b_l, c_l = [], [], []
    for (w0, delta, kappa), m in zip(poles, masks):
        denom = 1.0 + delta * dt
        a_l.append(np.where(m, dtype((2.0 - w0 ** 2 * dt ** 2) / denom), dtype(0.0)))
        b_l.append(np.where(m, dtype(-(1.0 - delta * dt) / denom), dtype(0.0)))
        c_l.append(np.where(m, dtype(EPS_0 * kappa * dt ** 2 / denom), dtype(0.0)))
    gamma = np.maximum(eps_inf + sig * dt_ / dtype(2.0), dtype(EPS_0 * 1e-10))
    ca = (eps_inf - sig * dt_ / dtype(2.0)) / gamma
    return ca, dt_ / gamma, a_l, b_l, c_l, dtype(1.0) / gamma


def debye_pole_masks(shape, coords, geometry, debye):
    """rfx/api/_compile.py:203-232, 352: per-pole masks, OR-ed over every
    shape whose material lists the pole, keyed by the pole's value in order
    of first appearance (equal poles of different materials merge), not
    replaced by later shapes and not extended into the CPML pads.
    debye: material name -> [pole tuple, ...] (Debye (delta_eps, tau) or
    Lorentz (omega_0, delta, kappa)).  Returns (poles, masks)."""
    poles, masks = [], []
    for lo, hi, name in geometry:
        for pole in debye.get(name, []):
            pole = tuple(float(v) for v in pole)
            if pole not in poles:
                poles.append(pole)
                masks.append(np.zeros(shape, bool))
            sel = np.ix_(*[np.nonzero(_axis_mask(coords[a], lo[a], hi[a]))[0]
                           for a in range(3)])
            masks[poles.index(pole)][sel] = True
    return poles, masks


def debye_coefficients(poles, masks, eps_r, sigma, dt, dtype=np.float64):
    """rfx/materials/debye.py:76-179 (init_debye): per pole
    alpha = (2 tau - dt)/(2 tau + dt), beta = eps0 d_eps dt/(2 tau + dt)
    inside its mask (0 elsewhere); gamma = eps0 eps_inf + sum beta +
    sigma dt/2 (floored at eps0 1e-10); ca = (eps0 eps_inf - sum beta -
    sigma dt/2)/gamma, cb = dt/gamma, cc_p = (1 - alpha_p)/gamma."""
    dt_ = dtype(dt)
    eps_inf = eps_r.astype(dtype) * dtype(EPS_0)
    sig = sigma.astype(dtype)
    alpha, beta = [], []
    for (de, tau), m in zip(poles, masks):
        a = (2.0 * tau - dt) / (2.0 * tau + dt)
        b = EPS_0 * de * dt / (2.0 * tau + dt)
        alpha.append(np.where(m, dtype(a), dtype(0.0)))
        beta.append(np.where(m, dtype(b), dtype(0.0)))
    beta_sum = np.zeros_like(eps_inf)
    for b in beta:
        beta_sum = beta_sum + b
    gamma = eps_inf + beta_sum + sig * dt_ / dtype(2.0)
    gamma = np.maximum(gamma, dtype(EPS_0 * 1e-10))
    ca = (eps_inf - beta_sum - sig * dt_ / dtype(2.0)) / gamma
    cb = dt_ / gamma
    cc = [(dtype(1.0) - a) / gamma for a in alpha]
    return ca, cb, alpha, beta, cc


def tangential_edge_masks(cell_mask):
    """rfx/boundaries/pec.py:90-176 (non-periodic axes: zero-padded shifts)."""
    out = []
    for ax in range(3):
        nb = shift_bwd(cell_mask, ax) | shift_fwd(cell_mask, ax)
        out.append(cell_mask & nb)
    return tuple(out)


# ---------------------------------------------------------------------------
# CPML profile (rfx/boundaries/cpml.py:101-171, 346-478)
# ---------------------------------------------------------------------------

def cpml_profile(n_layers, dt, d, kappa_max=1.0, order=3, R=1e-15):
    """rfx/boundaries/cpml.py:101-171 (float64 work, float32 result).

    Returns (b, c, kappa) float32 arrays for the lo face (index 0 =
    outermost cell).
    """
    eta = float(np.sqrt(MU_0 / EPS_0))
    thick = n_layers * d
    sigma_max = -float(np.log(R)) * (order + 1) / (2.0 * eta * thick)
    sigma_max = sigma_max * kappa_max
    rho = 1.0 - np.arange(n_layers, dtype=np.float64) / max(n_layers - 1, 1)
    sigma = sigma_max * rho ** order
    kappa = 1.0 + (kappa_max - 1.0) * rho ** order
    alpha = 0.05 * (1.0 - rho)
    denom = sigma * kappa + kappa ** 2 * alpha
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS_0)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(denom > 1e-30, sigma * (b - 1.0) / denom, 0.0)
    return b.astype(F32), c.astype(F32), kappa.astype(F32)


# ---------------------------------------------------------------------------
# Sources (rfx/sources/sources.py, rfx/simulation.py:181-315)
# ---------------------------------------------------------------------------

def gaussian_pulse_f32(times, f0, bandwidth=0.5, amplitude=1.0, cutoff=3.0):
    """rfx/sources/sources.py:56-104, evaluated in float32 like the JAX call.

    tau = 1/(f0 bw pi), t0 = cutoff tau, s = A (-2u) exp(-u^2), u=(t-t0)/tau.
    """
    tau = 1.0 / (f0 * bandwidth * math.pi)
    t0 = cutoff * tau
    t = np.asarray(times, dtype=F32)
    arg = (t - F32(t0)) / F32(tau)
    return (F32(amplitude) * (F32(-2.0) * arg) * np.exp(-(arg ** 2))).astype(F32)


def times_f32(n_steps, dt):
    """rfx/simulation.py:273: jnp.arange(n, float32) * dt -> float32."""
    return np.arange(n_steps, dtype=F32) * F32(dt)


def cell_cb_f32(dt, eps_r, sigma, idx):
    """rfx/simulation.py:181-194 in float32."""
    eps = eps_r[idx] * F32(EPS_0)
    sig = sigma[idx]
    loss = sig * F32(dt) / (F32(2.0) * eps)
    return (F32(dt) / eps) / (F32(1.0) + loss)


# ---------------------------------------------------------------------------
# Stencil helpers (rfx/core/yee.py:121-138)
# ---------------------------------------------------------------------------

def shift_fwd(a, ax):
    """a[i+1] with zero at the last index (yee.py:121-128)."""
    out = np.zeros_like(a)
    s_dst = [slice(None)] * 3
    s_src = [slice(None)] * 3
    s_dst[ax] = slice(0, -1)
    s_src[ax] = slice(1, None)
    out[tuple(s_dst)] = a[tuple(s_src)]
    return out


def shift_bwd(a, ax):
    """a[i-1] with zero at the first index (yee.py:131-138)."""
    out = np.zeros_like(a)
    s_dst = [slice(None)] * 3
    s_src = [slice(None)] * 3
    s_dst[ax] = slice(1, None)
    s_src[ax] = slice(0, -1)
    out[tuple(s_dst)] = a[tuple(s_src)]
    return out


def _bcast(v, ax):
    shp = [1, 1, 1]
    shp[ax] = -1
    return np.asarray(v).reshape(shp)


# ---------------------------------------------------------------------------
# The step (rfx/simulation.py:1256-1681; rfx/nonuniform.py:1256-1565)
# ---------------------------------------------------------------------------

# CPML correction table: (field, derivative-source field, sign) per axis.
# H: rfx/boundaries/cpml.py:875-1043; E: cpml.py:608-777.
CPML_H_TERMS = {
    0: (("hy", "ez", +1.0), ("hz", "ey", -1.0)),
    1: (("hx", "ez", -1.0), ("hz", "ex", +1.0)),
    2: (("hx", "ey", +1.0), ("hy", "ex", -1.0)),
}
CPML_E_TERMS = {
    0: (("ey", "hz", -1.0), ("ez", "hy", +1.0)),
    1: (("ex", "hz", +1.0), ("ez", "hx", -1.0)),
    2: (("ex", "hy", -1.0), ("ey", "hx", +1.0)),
}


# ---------------------------------------------------------------------------
# Subpixel smoothing: Kottke-averaged eps per E component
# (rfx/geometry/smoothing.py:454-611 uniform, :296-452 graded; box SDF
# :49-65, box normal :113-154, tensor average :244-293).  float32 throughout
# as the reference computes it with x64 off.
# ---------------------------------------------------------------------------
def _sdf_box_f32(X, Y, Z, lo, hi):
    lo = np.asarray(lo, F32)
    hi = np.asarray(hi, F32)
    c = (lo + hi) / F32(2.0)
    h = (hi - lo) / F32(2.0)
    dx = np.abs(X - c[0]) - h[0]
    dy = np.abs(Y - c[1]) - h[1]
    dz = np.abs(Z - c[2]) - h[2]
    ox, oy, oz = np.maximum(dx, F32(0)), np.maximum(dy, F32(0)), np.maximum(dz, F32(0))
    outside = np.sqrt(ox ** 2 + oy ** 2 + oz ** 2)
    inside = np.minimum(np.maximum(np.maximum(dx, dy), dz), F32(0))
    return (outside + inside).astype(F32)


def _normal_box_f32(X, Y, Z, lo, hi):
    lo = np.asarray(lo, F32)
    hi = np.asarray(hi, F32)
    c = (lo + hi) / F32(2.0)
    h = (hi - lo) / F32(2.0)
    rx = np.abs(X - c[0]) - h[0]
    ry = np.abs(Y - c[1]) - h[1]
    rz = np.abs(Z - c[2]) - h[2]
    is_x = (rx >= ry) & (rx >= rz)
    is_y = (ry > rx) & (ry >= rz)
    nx = np.where(is_x, np.sign(X - c[0]), F32(0)).astype(F32)
    ny = np.where(is_y, np.sign(Y - c[1]), F32(0)).astype(F32)
    nz = np.where(is_x | is_y, F32(0), np.sign(Z - c[2])).astype(F32)
    mag = np.sqrt(nx ** 2 + ny ** 2 + nz ** 2 + F32(1e-30))
    return nx / mag, ny / mag, nz / mag


def _kottke_f32(f, eps_in, eps_out, n):
    """Diagonal of the Kottke-averaged eps tensor (smoothing.py:244-293)."""
    e_in = F32(eps_in)
    inv_perp = f / e_in + (F32(1.0) - f) / (eps_out + F32(1e-30))
    eps_par = f * e_in + (F32(1.0) - f) * eps_out
    inv_par = F32(1.0) / (eps_par + F32(1e-30))
    out = []
    for nc in n:
        nn = nc * nc
        inv = nn * inv_perp + (F32(1.0) - nn) * inv_par
        out.append((F32(1.0) / (inv + F32(1e-30))).astype(F32))
    return out


def smoothed_eps(positions, length, shape, boxes, background=1.0):
    """Per-component eps_r of Kottke subpixel smoothing for axis-aligned
    boxes.  positions[comp] = (X, Y, Z) coordinate arrays (broadcastable to
    the grid) of that component's sample points, length = the cell length
    dividing the SDF (dx, or the graded lane's local geometric mean);
    boxes = [(lo, hi, eps_r)] in insertion order.  Boxes with the same eps_r
    form one group: union SDF (minimum), the normal of the box whose surface
    is nearest (first on ties); per group and component, cells with fill
    fraction f = clip(0.5 - sdf/length, 0, 1) of 1 take eps_r, cells with
    0 < f < 1 the Kottke average against the current value."""
    eps = [np.full(shape, F32(background), F32) for _ in range(3)]
    groups = {}
    for lo, hi, e in boxes:
        groups.setdefault(float(e), []).append((lo, hi))
    for e_r, members in groups.items():
        for comp in range(3):
            X, Y, Z = positions[comp]
            sdf = best = None
            nrm = None
            for lo, hi in members:
                sd = np.broadcast_to(_sdf_box_f32(X, Y, Z, lo, hi), shape)
                nv = [np.broadcast_to(v, shape) for v in _normal_box_f32(X, Y, Z, lo, hi)]
                if sdf is None:
                    sdf, best, nrm = sd, np.abs(sd), nv
                else:
                    sdf = np.minimum(sdf, sd)
                    closer = np.abs(sd) < best
                    best = np.where(closer, np.abs(sd), best)
                    nrm = [np.where(closer, a, b) for a, b in zip(nv, nrm)]
            f = np.clip(F32(0.5) - sdf / length, F32(0.0), F32(1.0)).astype(F32)
            inside = f >= F32(1.0)
            bnd = (f > F32(0.0)) & (f < F32(1.0))
            smooth = _kottke_f32(f, e_r, eps[comp], nrm)[comp]
            eps[comp] = np.where(inside, F32(e_r), np.where(bnd, smooth, eps[comp])).astype(F32)
    return tuple(eps)


def smoothed_eps_uniform(g, boxes):
    """smoothing.py:454-611: node coordinates f32(i - pad) * f32(dx), the
    component's axis shifted by half a cell, SDF over dx."""
    x, y, z = g.node_coords_f32()
    half = F32(g.dx * 0.5)
    X, Y, Z = x[:, None, None], y[None, :, None], z[None, None, :]
    pos = [(X + half, Y, Z), (X, Y + half, Z), (X, Y, Z + half)]
    return smoothed_eps(pos, F32(g.dx), g.shape, boxes)


def smoothed_eps_graded(g, boxes):
    """smoothing.py:296-452: centres node + d/2 along the component's axis,
    SDF over the local geometric mean (dx dy dz)^(1/3)."""
    nodes = g.node_coords_f32()
    d = [np.asarray(v, F32) for v in g.d32]
    cen = [nodes[a] + d[a] / F32(2.0) for a in range(3)]
    X, Y, Z = nodes[0][:, None, None], nodes[1][None, :, None], nodes[2][None, None, :]
    Xc, Yc, Zc = cen[0][:, None, None], cen[1][None, :, None], cen[2][None, None, :]
    length = (d[0][:, None, None] * d[1][None, :, None] * d[2][None, None, :]) ** F32(1.0 / 3.0)
    pos = [(Xc, Y, Z), (X, Yc, Z), (X, Y, Zc)]
    return smoothed_eps(pos, length.astype(F32), g.shape, boxes)


@dataclass
class StepConfig:
    """Everything the time loop needs, already set up on the host."""
    shape: tuple
    dt: float
    spacing: dict                     # {"kind": "uniform", "dx"} or graded inv arrays
    eps_r: np.ndarray                 # float32 (nx,ny,nz)
    sigma: np.ndarray
    mu_r: np.ndarray
    cpml_n: int = 0                   # layers per face (0 = no CPML)
    cpml_axes: str = "xyz"
    cpml_profiles: dict = field(default_factory=dict)  # (axis, side) -> (b,c,kappa) per slab index
    cpml_face_d: dict = field(default_factory=dict)    # axis -> (d_lo, d_hi)
    pec_axes: str = "xyz"
    pec_mask: np.ndarray | None = None
    sources: list = field(default_factory=list)        # (idx, comp, waveform f32 (n_steps,))
    probes: list = field(default_factory=list)         # (idx, comp)
    dft_planes: list = field(default_factory=list)     # (axis, index, comp, freqs)
    ports_vi: list = field(default_factory=list)       # (idx, comp, freqs, d_par, d_loop(dict))
    record_order: str = "uniform"                      # "uniform" | "graded"
    # flux monitors: (axis, index, lo1, hi1, lo2, hi2, freqs f32, weights f64 [n_steps+1] or None)
    flux: list = field(default_factory=list)
    # near-to-far-field boxes: (lo (i, j, k), hi (i, j, k), freqs f32)
    ntff: list = field(default_factory=list)
    # Debye poles [(delta_eps, tau)], Lorentz poles [(omega_0, delta, kappa)]
    # and their per-cell masks
    debye_poles: list = field(default_factory=list)
    debye_masks: list = field(default_factory=list)
    lorentz_poles: list = field(default_factory=list)
    lorentz_masks: list = field(default_factory=list)
    # subpixel smoothing: per-component eps_r (eps_ex, eps_ey, eps_ez) for
    # the E update's ca, cb (rfx/core/yee.py:824-901, graded :668-719)
    aniso_eps: tuple | None = None


class Stepper:
    """One FDTD time loop in the reference's sub-step order."""

    def __init__(self, cfg: StepConfig, dtype=np.float64):
        self.cfg = cfg
        self.dt = cfg.dt
        self.dtype = dtype
        dt = dtype(cfg.dt)
        eps = cfg.eps_r.astype(dtype) * dtype(EPS_0)
        sig = cfg.sigma.astype(dtype)
        mu = cfg.mu_r.astype(dtype) * dtype(MU_0)
        # yee.py:386-393 and :272/:289
        L = sig * dt / (dtype(2.0) * eps)
        self.ca = (dtype(1.0) - L) / (dtype(1.0) + L)
        self.cb = (dt / eps) / (dtype(1.0) + L)
        self.ch = dt / mu
        # cpml.py:538-546 (ce) and :817-826 (ch)
        self.ce_cpml = dt / (cfg.eps_r.astype(dtype) * dtype(EPS_0))
        self.ch_cpml = dt / (cfg.mu_r.astype(dtype) * dtype(MU_0))
        self.f = {c: np.zeros(cfg.shape, dtype) for c in COMPS}
        self.debye = self.lorentz = None
        self.aniso = None
        if cfg.aniso_eps is not None and not cfg.debye_poles and not cfg.lorentz_poles:
            # yee.py:880-895: per component, sigma isotropic
            self.aniso = []
            for e_c in cfg.aniso_eps:
                ea = e_c.astype(dtype) * dtype(EPS_0)
                La = sig * dt / (dtype(2.0) * ea)
                self.aniso.append(((dtype(1.0) - La) / (dtype(1.0) + La),
                                   (dt / ea) / (dtype(1.0) + La)))
        z3 = lambda: {c: np.zeros(cfg.shape, dtype) for c in ("ex", "ey", "ez")}  # noqa: E731
        if cfg.debye_poles:
            # with any pole every cell takes the dispersive ca, cb
            # (rfx/simulation.py:361-500)
            ca, cb, alpha, beta, cc = debye_coefficients(
                cfg.debye_poles, cfg.debye_masks, cfg.eps_r, cfg.sigma, cfg.dt, dtype)
            self.ca, self.cb = ca, cb
            self.debye = [alpha, beta, cc, [z3() for _ in cfg.debye_poles]]
        if cfg.lorentz_poles:
            ca_l, cb_l, a_l, b_l, c_l, cc_l = lorentz_coefficients(
                cfg.lorentz_poles, cfg.lorentz_masks, cfg.eps_r, cfg.sigma, cfg.dt, dtype)
            if self.debye is None:
                self.ca, self.cb = ca_l, cb_l
            else:
                # rfx/simulation.py:470-478
                alpha, beta, _, pol = self.debye
                beta_sum = np.zeros_like(ca_l)
                for b in beta:
                    beta_sum = beta_sum + b
                g_base = dtype(1.0) / cc_l
                g_tot = np.maximum(g_base + beta_sum, dtype(EPS_0 * 1e-10))
                self.ca = (ca_l * g_base - beta_sum) / g_tot
                self.cb = dtype(cfg.dt) / g_tot
                self.debye[2] = [(dtype(1.0) - a) / g_tot for a in alpha]
                cc_l = dtype(1.0) / g_tot
            self.lorentz = (a_l, b_l, c_l, cc_l, [z3() for _ in cfg.lorentz_poles],
                            [z3() for _ in cfg.lorentz_poles])
        self.psi = {}
        self.step_count = 0
        if cfg.pec_mask is not None:
            self.mask = tangential_edge_masks(cfg.pec_mask)
        else:
            self.mask = None
        sp = cfg.spacing
        if sp["kind"] == "graded":
            self.inv_e = [_bcast(v.astype(dtype), a) for a, v in enumerate(sp["inv_e"])]
            self.inv_h = [_bcast(v.astype(dtype), a) for a, v in enumerate(sp["inv_h"])]

    # -- differences ------------------------------------------------------
    def _dfwd(self, a, ax):
        sp = self.cfg.spacing
        if sp["kind"] == "uniform":
            return (shift_fwd(a, ax) - a) / self.dtype(sp["dx"])
        return (shift_fwd(a, ax) - a) * self.inv_h[ax]

    def _dbwd(self, a, ax):
        sp = self.cfg.spacing
        if sp["kind"] == "uniform":
            return (a - shift_bwd(a, ax)) / self.dtype(sp["dx"])
        return (a - shift_bwd(a, ax)) * self.inv_e[ax]

    # -- sub-steps ---------------------------------------------------------
    def update_h(self):
        """yee.py:265-291 (uniform) / yee.py:429-451 (graded)."""
        f = self.f
        cx = self._dfwd(f["ez"], 1) - self._dfwd(f["ey"], 2)
        cy = self._dfwd(f["ex"], 2) - self._dfwd(f["ez"], 0)
        cz = self._dfwd(f["ey"], 0) - self._dfwd(f["ex"], 1)
        f["hx"] = f["hx"] - self.ch * cx
        f["hy"] = f["hy"] - self.ch * cy
        f["hz"] = f["hz"] - self.ch * cz

    def update_e(self):
        """yee.py:379-404 (uniform) / yee.py:470-492 (graded)."""
        f = self.f
        cx = self._dbwd(f["hz"], 1) - self._dbwd(f["hy"], 2)
        cy = self._dbwd(f["hx"], 2) - self._dbwd(f["hz"], 0)
        cz = self._dbwd(f["hy"], 0) - self._dbwd(f["hx"], 1)
        if self.debye is None and self.lorentz is None:
            if self.aniso is not None:
                (ca0, cb0), (ca1, cb1), (ca2, cb2) = self.aniso
                f["ex"] = ca0 * f["ex"] + cb0 * cx
                f["ey"] = ca1 * f["ey"] + cb1 * cy
                f["ez"] = ca2 * f["ez"] + cb2 * cz
                return
            f["ex"] = self.ca * f["ex"] + self.cb * cx
            f["ey"] = self.ca * f["ey"] + self.cb * cy
            f["ez"] = self.ca * f["ez"] + self.cb * cz
            return
        # rfx/materials/lorentz.py:200-275, rfx/simulation.py:423-500,
        # rfx/materials/debye.py:182-250: Lorentz P first (explicit in E^n),
        # E^{n+1} = ca E + cb curl + sum cc_d P_d^n - cc_l sum (P_l^{n+1} - P_l^n),
        # then P_d^{n+1} = alpha P_d^n + beta (E^{n+1} + E^n)
        for comp, curl in (("ex", cx), ("ey", cy), ("ez", cz)):
            old = f[comp]
            new = self.ca * old + self.cb * curl
            if self.lorentz is not None:
                a_l, b_l, c_l, cc_l, pl, pl_prev = self.lorentz
                dp = np.zeros_like(new)
                for q in range(len(pl)):
                    pn = a_l[q] * pl[q][comp] + b_l[q] * pl_prev[q][comp] + c_l[q] * old
                    dp = dp + (pn - pl[q][comp])
                    pl_prev[q][comp] = pl[q][comp]
                    pl[q][comp] = pn
            if self.debye is not None:
                alpha, beta, cc, pol = self.debye
                acc = np.zeros_like(new)
                for q in range(len(pol)):
                    acc = acc + cc[q] * pol[q][comp]
                new = new + acc
            if self.lorentz is not None:
                new = new - cc_l * dp
            if self.debye is not None:
                for q in range(len(pol)):
                    pol[q][comp] = alpha[q] * pol[q][comp] + beta[q] * (new + old)
            f[comp] = new

    def _slab(self, ax, side):
        n = self.cfg.cpml_n
        s = [slice(None)] * 3
        N = self.cfg.shape[ax]
        s[ax] = slice(0, n) if side == 0 else slice(N - n, N)
        return tuple(s)

    def _cpml(self, terms, is_h):
        cfg = self.cfg
        if cfg.cpml_n <= 0:
            return
        dt_ = self.dtype
        coef = self.ch_cpml if is_h else self.ce_cpml
        for ax in range(3):
            name = "xyz"[ax]
            if name not in cfg.cpml_axes:
                continue
            for side in (0, 1):
                sl = self._slab(ax, side)
                b, c, k = cfg.cpml_profiles[(ax, side)]
                b = _bcast(b.astype(dt_), ax)
                c = _bcast(c.astype(dt_), ax)
                kk = _bcast(k.astype(dt_), ax)
                d = dt_(cfg.cpml_face_d[ax][side])
                for fld, src, sign in terms[ax]:
                    a = self.f[src]
                    if is_h:
                        g = (shift_fwd(a, ax)[sl] - a[sl]) / d
                    else:
                        g = (a[sl] - shift_bwd(a, ax)[sl]) / d
                    key = (fld, ax, side)
                    psi = self.psi.get(key)
                    if psi is None:
                        psi = np.zeros_like(g)
                    psi = b * psi + c * g
                    self.psi[key] = psi
                    cc = coef[sl]
                    upd = self.f[fld][sl]
                    upd = upd + dt_(sign) * cc * psi
                    upd = upd + dt_(sign) * cc * (dt_(1.0) / kk - dt_(1.0)) * g
                    self.f[fld][sl] = upd

    def apply_pec(self):
        """rfx/boundaries/pec.py:13-48."""
        f = self.f
        ax = self.cfg.pec_axes
        if "x" in ax:
            for c in ("ey", "ez"):
                f[c][0, :, :] = 0.0
                f[c][-1, :, :] = 0.0
        if "y" in ax:
            for c in ("ex", "ez"):
                f[c][:, 0, :] = 0.0
                f[c][:, -1, :] = 0.0
        if "z" in ax:
            for c in ("ex", "ey"):
                f[c][:, :, 0] = 0.0
                f[c][:, :, -1] = 0.0
            f["ez"][:, :, -1] = 0.0
        if self.mask is not None:  # pec.py:179-208
            for c, m in zip(("ex", "ey", "ez"), self.mask):
                f[c] = f[c] * (1.0 - m.astype(self.dtype))

    def step_fields(self):
        """H -> CPML-H -> E -> CPML-E -> PEC (simulation.py:1256-1407)."""
        self.update_h()
        self._cpml(CPML_H_TERMS, True)
        self.update_e()
        self._cpml(CPML_E_TERMS, False)
        self.apply_pec()
        self.step_count += 1


def port_current(f, idx, comp, loop):
    """Ampere loop of H around the port edge (simulation.py:1484-1493).

    ``loop`` is the per-axis length multiplying each H difference: uniform
    lane dx (simulation.py:1484-1493); graded wire ports use the dual
    spacings (rfx/nonuniform.py:720-777).
    """
    i, j, k = idx
    if comp == "ez":
        return ((f["hy"][i, j, k] - f["hy"][i - 1, j, k]) * loop[1]
                - (f["hx"][i, j, k] - f["hx"][i, j - 1, k]) * loop[0])
    if comp == "ex":
        return ((f["hz"][i, j, k] - f["hz"][i, j - 1, k]) * loop[2]
                - (f["hy"][i, j, k] - f["hy"][i, j, k - 1]) * loop[1])
    return ((f["hx"][i, j, k] - f["hx"][i, j, k - 1]) * loop[0]
            - (f["hz"][i, j, k] - f["hz"][i - 1, j, k]) * loop[2])


@dataclass
class RunOutput:
    time_series: np.ndarray
    dft_planes: list
    port_v: list        # per port: (n_steps,) V samples (pre-injection)
    port_i: list
    port_v_dft: list
    port_i_dft: list
    fields: dict
    psi: dict
    flux: list = field(default_factory=list)   # per monitor (e1, e2, h1, h2) complex128 DFTs
    ntff: list = field(default_factory=list)   # per box {face: (n_freqs, n1, n2, 4) complex128}


def run_loop(cfg: StepConfig, n_steps: int, dtype=np.float64, stepper=None):
    """rfx/simulation.py:1977 scan body order, per step n:
    fields -> port V/I sample (t = n dt, before injection, :1449-1497) ->
    sources (:1520-1523) -> probes (:1565-1567) -> DFT planes
    (t = (n+1) dt, :1574-1588).  Graded order (rfx/nonuniform.py:1354-1542):
    sources -> port V/I -> DFT planes (t = n dt) -> probes.

    DFT sums are accumulated in float64 (complex128).
    """
    st = stepper if stepper is not None else Stepper(cfg, dtype)
    f = st.f
    npr = len(cfg.probes)
    ts = np.zeros((n_steps, npr), dtype=np.float64)
    accs = [np.zeros((len(fr),) + _plane_shape(cfg.shape, ax), np.complex128)
            for ax, _, _, fr in cfg.dft_planes]
    pv = [np.zeros(n_steps) for _ in cfg.ports_vi]
    pi = [np.zeros(n_steps) for _ in cfg.ports_vi]
    fl = [tuple(np.zeros((len(m[6]), m[3] - m[2], m[5] - m[4]), np.complex128) for _ in range(4))
          for m in cfg.flux]
    nt = [{name: np.zeros((len(fr),) + shp + (4,), np.complex128)
           for name, shp in zip(NTFF_FACES, ntff_face_shapes(lo, hi))} for lo, hi, fr in cfg.ntff]
    dt = cfg.dt
    for n in range(n_steps):
        st.step_fields()
        if cfg.record_order == "uniform":
            _sample_ports(f, cfg, pv, pi, n)
            _inject(f, cfg, n)
            _probe(f, cfg, ts, n)
            _planes(f, cfg, accs, (n + 1) * dt, dt)
            _flux(f, cfg, fl, n + 1, dt)
        else:
            _inject(f, cfg, n)
            _sample_ports(f, cfg, pv, pi, n)
            _planes(f, cfg, accs, n * dt, dt)
            _flux(f, cfg, fl, n, dt)
            _probe(f, cfg, ts, n)
        _ntff(f, cfg, nt, n, dt)
    tvec = np.arange(n_steps) * dt
    vd, idft = [], []
    for p, (idx, comp, freqs, d_par, loop) in enumerate(cfg.ports_vi):
        ph = np.exp(-1j * 2.0 * np.pi * np.outer(np.asarray(freqs, np.float64), tvec)) * dt
        vd.append(ph @ pv[p])
        idft.append(ph @ pi[p])
    return RunOutput(ts, accs, pv, pi, vd, idft, f, st.psi, fl, nt)


def _plane_shape(shape, ax):
    return tuple(s for a, s in enumerate(shape) if a != ax)


def _plane(f, ax, index):
    if ax == 0:
        return f[index, :, :]
    if ax == 1:
        return f[:, index, :]
    return f[:, :, index]


def _sample_ports(f, cfg, pv, pi, n):
    for p, (idx, comp, freqs, d_par, loop) in enumerate(cfg.ports_vi):
        pv[p][n] = -f[comp][idx] * d_par
        pi[p][n] = port_current(f, idx, comp, loop)


def _inject(f, cfg, n):
    for idx, comp, wave in cfg.sources:
        f[comp][idx] += wave[n]


def _probe(f, cfg, ts, n):
    for p, (idx, comp) in enumerate(cfg.probes):
        ts[n, p] = f[comp][idx]


def _planes(f, cfg, accs, t, dt):
    for q, (ax, index, comp, freqs) in enumerate(cfg.dft_planes):
        ph = np.exp(-1j * 2.0 * np.pi * np.asarray(freqs, np.float64) * t) * dt
        accs[q] += ph[:, None, None] * _plane(f[comp], ax, index)[None].astype(np.float64)


# rfx/probes/probes.py:506-510: tangential (E1, E2, H1, H2) per plane normal
FLUX_COMPONENTS = {0: ("ey", "ez", "hy", "hz"), 1: ("ez", "ex", "hz", "hx"),
                   2: ("ex", "ey", "hx", "hy")}


def _flux(f, cfg, fl, step, dt):
    """Flux-monitor DFTs (rfx/simulation.py:1590-1644 uniform, step = n+1;
    rfx/nonuniform.py:1484-1522 graded, step = n): E of the plane at
    t = step dt, H averaged over the planes index-1 and index (co-located
    with E) at t - dt/2, times dt and the window weight of `step`."""
    for q, (ax, idx, lo1, hi1, lo2, hi2, freqs, wts) in enumerate(cfg.flux):
        e1n, e2n, h1n, h2n = FLUX_COMPONENTS[ax]
        im1 = max(idx - 1, 0)

        def sl(a, i):
            p = _plane(a, ax, i)
            return p[lo1:hi1, lo2:hi2].astype(np.float64)

        w = 1.0 if wts is None else float(wts[step])
        t = step * dt
        fr = np.asarray(freqs, np.float64)
        ke = np.exp(-1j * 2.0 * np.pi * fr * t) * dt * w
        kh = np.exp(-1j * 2.0 * np.pi * fr * (t - dt * 0.5)) * dt * w
        hv1 = ((_plane(f[h1n], ax, im1) + _plane(f[h1n], ax, idx)) * 0.5)[lo1:hi1, lo2:hi2]
        hv2 = ((_plane(f[h2n], ax, im1) + _plane(f[h2n], ax, idx)) * 0.5)[lo1:hi1, lo2:hi2]
        e1, e2, a1, a2 = fl[q]
        e1 += ke[:, None, None] * sl(f[e1n], idx)[None]
        e2 += ke[:, None, None] * sl(f[e2n], idx)[None]
        a1 += kh[:, None, None] * hv1.astype(np.float64)[None]
        a2 += kh[:, None, None] * hv2.astype(np.float64)[None]


# rfx/farfield.py:81-84: tangential components stored per face normal
NTFF_COMPONENTS = {0: ("ey", "ez", "hy", "hz"), 1: ("ex", "ez", "hx", "hz"),
                   2: ("ex", "ey", "hx", "hy")}
NTFF_FACES = ("x_lo", "x_hi", "y_lo", "y_hi", "z_lo", "z_hi")


def ntff_face_shapes(lo, hi):
    ni, nj, nk = (hi[a] - lo[a] for a in range(3))
    return [(nj, nk), (nj, nk), (ni, nk), (ni, nk), (ni, nj), (ni, nj)]


def _ntff(f, cfg, nt, n, dt):
    """rfx/farfield.py:208-313 (accumulate_ntff) with float64 sums: per
    face the tangential (E1, E2, H1, H2) of the plane itself over the box's
    half-open ranges; E at t = n dt, H at t = n dt + dt/2 (n = scan index,
    on both lanes: rfx/simulation.py:1570-1572, rfx/nonuniform.py:1527)."""
    for q, (lo, hi, freqs) in enumerate(cfg.ntff):
        fr = np.asarray(freqs, np.float64)
        t = n * dt
        pe = np.exp(-1j * 2.0 * np.pi * fr * t) * dt
        ph = np.exp(-1j * 2.0 * np.pi * fr * (t + dt * 0.5)) * dt
        kern = np.stack([pe, pe, ph, ph], axis=-1)[:, None, None, :]
        rng = [slice(lo[a], hi[a]) for a in range(3)]
        for face, name in enumerate(NTFF_FACES):
            ax = face // 2
            idx = hi[ax] if face & 1 else lo[ax]
            sl = list(rng)
            sl[ax] = idx
            vals = np.stack([f[c][tuple(sl)].astype(np.float64) for c in NTFF_COMPONENTS[ax]],
                            axis=-1)
            nt[q][name] += kern * vals[None]


def ntff_axis_tables(shape, dx, dy=None, dz=None, cpml_lo=(0, 0, 0)):
    """Face positions and widths per axis (rfx/farfield.py:344-390 and
    :456-462): (index - cpml_lo) * d on a uniform axis, the cumulative z
    edges (float64 sums of the float32 dz profile, origin at cpml_lo) on a
    graded z axis; widths d, or dz[k]."""
    dy = dx if dy is None else dy
    pos = [(np.arange(shape[0]) - cpml_lo[0]) * float(dx),
           (np.arange(shape[1]) - cpml_lo[1]) * float(dy)]
    width = [np.full(shape[0], float(dx)), np.full(shape[1], float(dy))]
    if dz is None:
        pos.append((np.arange(shape[2]) - cpml_lo[2]) * float(dx))
        width.append(np.full(shape[2], float(dx)))
    else:
        d64 = np.asarray(dz, np.float64)
        edges = np.concatenate([[0.0], np.cumsum(d64)])
        pos.append((edges - edges[cpml_lo[2]])[:shape[2]])
        width.append(d64)
    return pos, width


def far_field(faces, lo, hi, freqs, pos, width, theta, phi, c0=C0, eta0=None):
    """rfx/farfield.py:393-556 (compute_far_field) in float64: surface
    currents J = n x H, M = -n x E per face (outward normal n = -/+ axis on
    the lo / hi face), N = sum J e^{j k r.p} dS, L = sum M e^{j k r.p} dS,
    E_theta = -j k/(4 pi) (L_phi + eta0 N_theta),
    E_phi = j k/(4 pi) (L_theta - eta0 N_phi).  Returns (E_theta, E_phi),
    each (n_freqs, n_theta, n_phi)."""
    if eta0 is None:
        eta0 = float(np.sqrt(MU_0 / EPS_0))
    theta = np.asarray(theta, np.float64)
    phi = np.asarray(phi, np.float64)
    fr = np.asarray(freqs, np.float64)
    k = 2 * np.pi * fr / c0
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    st, ct, sp, cp = np.sin(TH), np.cos(TH), np.sin(PH), np.cos(PH)
    r_hat = np.stack([st * cp, st * sp, ct], -1).reshape(-1, 3)
    th_hat = np.stack([ct * cp, ct * sp, -st], -1).reshape(-1, 3)
    ph_hat = np.stack([-sp, cp, np.zeros_like(st)], -1).reshape(-1, 3)
    N = np.zeros((len(fr), r_hat.shape[0], 3), np.complex128)
    L = np.zeros_like(N)
    for face, name in enumerate(NTFF_FACES):
        ax = face // 2
        normal = np.zeros(3)
        normal[ax] = 1.0 if face & 1 else -1.0
        t1, t2 = [a for a in range(3) if a != ax]
        data = np.asarray(faces[name], np.complex128)           # (nf, n1, n2, 4)
        i1 = np.arange(lo[t1], hi[t1])
        i2 = np.arange(lo[t2], hi[t2])
        P = np.zeros((len(i1), len(i2), 3))
        P[..., ax] = pos[ax][hi[ax] if face & 1 else lo[ax]]
        P[..., t1] = pos[t1][i1][:, None]
        P[..., t2] = pos[t2][i2][None, :]
        dS = width[t1][i1][:, None] * width[t2][i2][None, :]
        # tangential E and H as 3-vectors, then the cross products
        Ev = np.zeros(data.shape[:3] + (3,), np.complex128)
        Hv = np.zeros_like(Ev)
        Ev[..., t1], Ev[..., t2] = data[..., 0], data[..., 1]
        Hv[..., t1], Hv[..., t2] = data[..., 2], data[..., 3]
        J = np.cross(normal, Hv)
        M = -np.cross(normal, Ev)
        dot = P.reshape(-1, 3) @ r_hat.T                        # (nc, n_dir)
        for fi in range(len(fr)):
            e = np.exp(1j * k[fi] * dot)
            N[fi] += e.T @ (J[fi] * dS[..., None]).reshape(-1, 3)
            L[fi] += e.T @ (M[fi] * dS[..., None]).reshape(-1, 3)
    n_th = np.einsum("fdc,dc->fd", N, th_hat)
    n_ph = np.einsum("fdc,dc->fd", N, ph_hat)
    l_th = np.einsum("fdc,dc->fd", L, th_hat)
    l_ph = np.einsum("fdc,dc->fd", L, ph_hat)
    g = (1j * k / (4 * np.pi))[:, None]
    e_th = -g * (l_ph + eta0 * n_th)
    e_ph = g * (l_th - eta0 * n_ph)
    shp = (len(fr), len(theta), len(phi))
    return e_th.reshape(shp), e_ph.reshape(shp)


def dft_window_weights(total_steps, window="rect", alpha=0.25):
    """rfx/core/dft_utils.py:8-49 in float32, for steps 0..total_steps."""
    if total_steps <= 1 or window == "rect":
        return None
    n = np.arange(total_steps + 1, dtype=F32)
    N = F32(total_steps - 1)
    x = n / np.maximum(N, F32(1.0))
    two_pi = F32(2.0 * np.pi)
    if window == "hann" or (window == "tukey" and alpha >= 1.0):
        w = F32(0.5) * (F32(1.0) - np.cos(two_pi * x))
    elif window == "tukey":
        a = F32(alpha)
        if alpha <= 0.0:
            return None
        left = F32(0.5) * (F32(1.0) + np.cos(F32(np.pi) * (F32(2.0) * x / a - F32(1.0))))
        right = F32(0.5) * (F32(1.0) + np.cos(F32(np.pi) * (F32(2.0) * x / a - F32(2.0) / a
                                                             + F32(1.0))))
        w = np.where(x < a / F32(2.0), left, np.where(x <= F32(1.0) - a / F32(2.0), F32(1.0),
                                                       right))
    else:
        raise ValueError(f"unknown DFT window {window!r}")
    return w.astype(np.float64)


def flux_spectrum(accs, dA):
    """rfx/probes/probes.py:615-660: sum Re(E1 H2* - E2 H1*) dA (float64)."""
    e1, e2, h1, h2 = accs
    return np.real(np.sum((e1 * np.conj(h2) - e2 * np.conj(h1)) * np.asarray(dA, np.float64),
                          axis=(-2, -1)))


def extract_lumped_s11(v_dft, i_dft, z0=50.0):
    """rfx/probes/probes.py:1239-1273: S11 = (V + Z0 I) / (V - Z0 I)."""
    den = v_dft - z0 * i_dft
    den = np.where(np.abs(den) > 0.0, den, 1.0)
    return (v_dft + z0 * i_dft) / den


# ---------------------------------------------------------------------------
# High-level uniform-lane set-up (rfx/api/_compile.py, rfx/runners/uniform.py,
# rfx/api/_execute.py:917-1669)
# ---------------------------------------------------------------------------

@dataclass
class Pulse:
    f0: float
    bandwidth: float = 0.5
    amplitude: float = 1.0
    cutoff: float = 3.0


def _smoothing_boxes(spec):
    mats = spec.get("materials", {})
    return [(lo, hi, mats[name][0]) for lo, hi, name in spec.get("geometry", [])]


def build_uniform(spec, n_steps):
    """Build a StepConfig for the uniform lane from a high-level spec.

    spec keys: freq_max, domain, dx, boundary ("cpml"|"pec"), cpml_layers,
    kappa_max, materials {name: (eps_r, sigma, mu_r)}, geometry
    [(lo, hi, name)], sources [(pos, comp, Pulse)], ports
    [(pos, comp, z0, Pulse|None, excite)], probes [(pos, comp)],
    dft_planes [(axis, coord, comp, freqs|None, n_freqs)], lane
    ("run"|"forward"), eps_override, sigma_override, mu_r_override (arrays
    or None; forward lane: each replaces its array wholesale,
    rfx/api/_execute.py:2800-2805), s11_freqs.

    Lane semantics: ``run`` applies outer-face PEC on every axis
    (rfx/runners/uniform.py:321 passes pec_axes=None -> all axes,
    rfx/simulation.py:684-691); ``forward`` passes pec_axes = axes without
    CPML, i.e. none when CPML covers x, y and z (rfx/api/_execute.py:
    1420-1421).  Soft sources are Cb-normalised J sources on CPML
    (uniform.py:339-360) and always on forward() (_execute.py:1040-1051);
    raw field adds on the run() PEC lane (simulation.py:205-236).
    """
    boundary = spec.get("boundary", "cpml")
    layers = spec.get("cpml_layers", 16) if boundary == "cpml" else 0
    g = UniformGrid(spec["freq_max"], spec["domain"], spec.get("dx"), layers)
    shape = g.shape
    coords = g.node_coords_f32()
    eps_r, sigma, mu_r, pec = assemble_materials(
        shape, coords, spec.get("geometry", []), spec.get("materials", {}),
        layers, spec.get("material_sweep"))
    lane = spec.get("lane", "run")
    if lane == "forward" and spec.get("eps_override") is not None:
        # _execute.py:2800-2805: override replaces eps_r wholesale
        eps_r = np.asarray(spec["eps_override"], F32).reshape(shape).copy()
    if lane == "forward" and spec.get("sigma_override") is not None:
        sigma = np.asarray(spec["sigma_override"], F32).reshape(shape).copy()
    if lane == "forward" and spec.get("mu_r_override") is not None:
        mu_r = np.asarray(spec["mu_r_override"], F32).reshape(shape).copy()
    if lane == "forward" and spec.get("pec_mask_override") is not None:
        # _execute.py:2807-2808: merged with the geometry's PEC cells; the
        # ports below clear their own cells as for geometry PEC
        extra = np.asarray(spec["pec_mask_override"], bool).reshape(shape)
        pec = extra.copy() if pec is None else (pec | extra)
    dt = g.dt
    tms = times_f32(n_steps, dt)
    sources, ports_vi = [], []
    s11_freqs = spec.get("s11_freqs")
    # add_source() and add_port() both append to Simulation._ports in
    # registration order (rfx/api/__init__.py:1008-1016, 1609-1624); a soft
    # source is the impedance-0 entry.  Sources are built in that order.
    for pos, comp, z0, pulse, excite in spec.get("ports", []):
        idx = g.position_to_index(pos)
        if pulse is None and excite:
            pulse = Pulse(spec["freq_max"] / 2, 0.8)
        if z0 == 0.0:
            # simulation.py:239-293 (J source) / :205-236 (raw add)
            wave = gaussian_pulse_f32(tms, pulse.f0, pulse.bandwidth,
                                      pulse.amplitude, pulse.cutoff)
            if boundary == "cpml" or lane == "forward":
                wave = cell_cb_f32(dt, eps_r, sigma, idx) * wave
            sources.append((idx, comp, wave.astype(F32)))
            continue
        # sources.py:31-45 (port_sigma) + :238-247 (fold into sigma)
        sp = g.dx / (z0 * g.dx * g.dx)
        sigma = sigma.copy()
        sigma[idx] = F32(sigma[idx] + F32(sp))
        if pec is not None:
            pec = pec.copy()
            pec[idx] = False
        if excite:
            # simulation.py:296-315
            cb = cell_cb_f32(dt, eps_r, sigma, idx)
            w = (cb / F32(g.dx)) * gaussian_pulse_f32(
                tms, pulse.f0, pulse.bandwidth, pulse.amplitude, pulse.cutoff)
            sources.append((idx, comp, w.astype(F32)))
        if s11_freqs is not None:
            ports_vi.append((idx, comp, np.asarray(s11_freqs, F32),
                             g.dx, (g.dx, g.dx, g.dx)))
    probes = [(g.position_to_index(p), c) for p, c in spec.get("probes", [])]
    if not probes and lane == "forward" and spec.get("ports"):
        # _execute.py:1319-1321: no probes -> one probe per port entry
        probes = [(g.position_to_index(p[0]), p[1]) for p in spec["ports"]]
    planes = []
    for axis, coord, comp, freqs, nf in spec.get("dft_planes", []):
        a = "xyz".index(axis)
        pos = [0.0, 0.0, 0.0]
        pos[a] = coord
        index = g.position_to_index(tuple(pos))[a]
        if freqs is None:
            freqs = np.linspace(spec["freq_max"] / 10, spec["freq_max"], nf).astype(F32)
        planes.append((a, index, comp, np.asarray(freqs, F32)))
    profiles, face_d = {}, {}
    if layers > 0:
        b, c, k = cpml_profile(layers, dt, g.dx, spec.get("kappa_max", 1.0))
        for ax in range(3):
            profiles[(ax, 0)] = (b, c, k)
            profiles[(ax, 1)] = (b[::-1].copy(), c[::-1].copy(), k[::-1].copy())
            face_d[ax] = (g.dx, g.dx)
    # forward(): pec_axes = axes outside grid.cpml_axes, and Grid keeps
    # cpml_axes = "xyz" even with zero layers (rfx/grid.py:86), so forward()
    # applies no outer-face PEC on either boundary kind.
    pec_axes = "xyz" if lane == "run" else ""
    flux, areas = _flux_entries(spec, g, n_steps, False)
    cfg = StepConfig(shape=shape, dt=dt, spacing=g.spacing(), eps_r=eps_r,
                     sigma=sigma, mu_r=mu_r, cpml_n=layers,
                     cpml_axes="xyz" if layers > 0 else "",
                     cpml_profiles=profiles, cpml_face_d=face_d,
                     pec_axes=pec_axes, pec_mask=pec, sources=sources,
                     probes=probes, dft_planes=planes, ports_vi=ports_vi, flux=flux)
    cfg.flux_area = areas
    cfg.ntff = _ntff_entries(spec, g)
    if spec.get("subpixel") and lane == "run":
        # rfx/runners/uniform.py:271-278: every geometry entry with its
        # material's eps_r, background 1
        cfg.aniso_eps = smoothed_eps_uniform(g, _smoothing_boxes(spec))
    if spec.get("debye"):
        cfg.debye_poles, cfg.debye_masks = debye_pole_masks(
            shape, coords, spec.get("geometry", []), spec["debye"])
    if spec.get("lorentz"):
        cfg.lorentz_poles, cfg.lorentz_masks = debye_pole_masks(
            shape, coords, spec.get("geometry", []), spec["lorentz"])
    return g, cfg


def simulate_uniform(spec, n_steps, dtype=np.float64):
    """Set up and run the uniform lane; returns (grid, cfg, RunOutput)."""
    g, cfg = build_uniform(spec, n_steps)
    out = run_loop(cfg, n_steps, dtype)
    return g, cfg, out


# ---------------------------------------------------------------------------
# Graded-z lane set-up (rfx/runners/nonuniform.py:465-925, rfx/nonuniform.py)
# ---------------------------------------------------------------------------

def current_source_f32(g: GradedGrid, eps_r, sigma, idx, comp, dt, wave_f32):
    """rfx/nonuniform.py:588-691 make_current_source: w = f32(Cb / dV) * s(t),
    Cb from the cell's (float64-promoted) eps_r and sigma, dV = primal width
    on the component's axis x dual spacings on the two transverse axes."""
    eps = float(eps_r[idx]) * EPS_0
    sig = float(sigma[idx])
    loss = sig * dt / (2.0 * eps)
    cb = (dt / eps) / (1.0 + loss)
    a = {"ex": 0, "ey": 1, "ez": 2}[comp]
    w = []
    for ax in range(3):
        if ax == a:
            w.append(float(g.d32[ax][idx[ax]]))
        else:
            w.append(g.dual64(ax, idx[ax]))
    dV = w[0] * w[1] * w[2]
    return (F32(cb / dV) * wave_f32).astype(F32)


def _graded_port_metrics(g: GradedGrid, cell, axis):
    """d_parallel (primal) and the two transverse dual spacings of a port
    cell (rfx/runners/nonuniform.py:766-796, 846-868)."""
    ci, cj, ck = cell
    dxi = float(g.d32[0][ci])
    dyj = float(g.d32[1][cj])
    dual_x = g.dual64(0, ci)
    dual_y = g.dual64(1, cj)
    dual_z = g.dual_z32(ck)
    if axis == 2:
        return float(g.d32[2][ck]), dual_x, dual_y, (dual_x, dual_y, dual_z)
    if axis == 1:
        return dyj, dual_x, dual_z, (dual_x, dual_y, dual_z)
    return dxi, dual_y, dual_z, (dual_x, dual_y, dual_z)


def build_graded(spec, n_steps):
    """StepConfig for the graded-z lane (Simulation(dz_profile=...).run()).

    spec keys as build_uniform plus ``dz_profile``; ``domain`` is (Lx, Ly)
    (z comes from the profile); a port entry may carry a sixth element
    ``extent`` (wire port).  Order of set-up (rfx/runners/nonuniform.py):
    grid (build_nonuniform_grid :25-60, dx default C0/freq_max/20),
    materials on node coordinates + CPML pad extension
    (assemble_materials_nu :63-140), optional eps/sigma override
    (:630-643), ports in registration order (:694-887): soft sources are
    current sources normalised by the CONCRETE (pre-override, pre-fold)
    materials; a single-cell port folds sigma = d_par/(Z0 dual1 dual2) and
    injects a current source; a wire port folds n_live d/(Z0 dual1 dual2)
    on every live cell and injects source/n_live on each, and records V/I
    at its middle cell.  Outer-face PEC always (rfx/nonuniform.py:1319);
    CPML kappa_max 1 (init_cpml reads the grid's kappa_max, which the
    graded grid does not carry, rfx/boundaries/cpml.py:391-392).
    """
    boundary = spec.get("boundary", "cpml")
    layers = spec.get("cpml_layers", 16) if boundary == "cpml" else 0
    dx = spec.get("dx")
    if dx is None:
        dx = C0 / spec["freq_max"] / 20.0
    dz = spec.get("dz_profile")
    if dz is None:
        # rfx/runners/nonuniform.py:47-55: uniform z from the domain
        dz = np.full(max(1, int(round(spec["domain"][2] / dx))), float(dx))
    g = GradedGrid(dz, dx, spec["domain"][:2], layers, dx_profile=spec.get("dx_profile"),
                   dy_profile=spec.get("dy_profile"))
    shape = g.shape
    coords = g.node_coords_f32()
    eps_r, sigma, mu_r, pec = assemble_materials(
        shape, coords, spec.get("geometry", []), spec.get("materials", {}), layers)
    eps_c, sig_c = eps_r.copy(), sigma.copy()
    if spec.get("eps_override") is not None:
        eps_r = np.asarray(spec["eps_override"], F32).reshape(shape).copy()
    if spec.get("sigma_override") is not None:
        sigma = np.asarray(spec["sigma_override"], F32).reshape(shape).copy()
    dt = g.dt
    tms = times_f32(n_steps, dt)
    sources, ports_vi = [], []
    s11_freqs = spec.get("s11_freqs")
    for entry in spec.get("ports", []):
        pos, comp, z0, pulse, excite = entry[:5]
        extent = entry[5] if len(entry) > 5 else None
        idx = g.position_to_index(pos)
        if pulse is None and excite:
            pulse = Pulse(spec["freq_max"] / 2, 0.8)
        wave = None
        if pulse is not None:
            wave = gaussian_pulse_f32(tms, pulse.f0, pulse.bandwidth, pulse.amplitude,
                                      pulse.cutoff)
        axis = {"ex": 0, "ey": 1, "ez": 2}[comp]
        if z0 == 0.0:
            sources.append((idx, comp, current_source_f32(g, eps_c, sig_c, idx, comp, dt, wave)))
            continue
        if extent is None:
            d_par, p1, p2, _ = _graded_port_metrics(g, idx, axis)
            sigma = sigma.copy()
            sigma[idx] = F32(sigma[idx] + F32(d_par / (z0 * p1 * p2)))
            if pec is not None:
                pec = pec.copy()
                pec[idx] = False
            if excite:
                sources.append((idx, comp, current_source_f32(g, eps_c, sig_c, idx, comp, dt,
                                                              wave)))
            continue
        end = list(pos)
        end[axis] += extent
        idx_end = g.position_to_index(tuple(end))
        lo, hi = sorted((idx[axis], idx_end[axis]))
        cells = []
        for k in range(lo, hi + 1):
            c = list(idx)
            c[axis] = k
            cells.append(tuple(c))
        live = [True if pec is None else not bool(pec[c]) for c in cells]
        n_live = sum(live)
        if n_live == 0:
            raise ValueError("wire port: every extent cell lies inside PEC")
        sigma = sigma.copy()
        for c, lv in zip(cells, live):
            if not lv:
                continue
            d_par, p1, p2, _ = _graded_port_metrics(g, c, axis)
            sigma[c] = F32(sigma[c] + F32(n_live * d_par / (z0 * p1 * p2)))
            if pec is not None:
                pec = pec.copy()
                pec[c] = False
        if excite:
            for c, lv in zip(cells, live):
                if lv:
                    w = current_source_f32(g, eps_c, sig_c, c, comp, dt, wave)
                    sources.append((c, comp, (w / n_live).astype(F32)))
        mid = cells[len(cells) // 2]
        if s11_freqs is not None:
            d_par, _, _, loop = _graded_port_metrics(g, mid, axis)
            ports_vi.append((mid, comp, np.asarray(s11_freqs, F32), d_par, loop))
    probes = [(g.position_to_index(p), c) for p, c in spec.get("probes", [])]
    planes = []
    for axis_name, coord, comp, freqs, nf in spec.get("dft_planes", []):
        a = "xyz".index(axis_name)
        pos = [0.0, 0.0, 0.0]
        pos[a] = coord
        index = g.position_to_index(tuple(pos))[a]
        if freqs is None:
            freqs = np.linspace(spec["freq_max"] / 10, spec["freq_max"], nf).astype(F32)
        planes.append((a, index, comp, np.asarray(freqs, F32)))
    profiles, face_d = {}, {}
    if layers > 0:
        for ax, name in enumerate("xyz"):
            d_lo, d_hi = g.face_d[name]
            b, c, k = cpml_profile(layers, dt, d_lo, 1.0)
            profiles[(ax, 0)] = (b, c, k)
            b, c, k = cpml_profile(layers, dt, d_hi, 1.0)
            profiles[(ax, 1)] = (b[::-1].copy(), c[::-1].copy(), k[::-1].copy())
            face_d[ax] = (d_lo, d_hi)
    cfg = StepConfig(shape=shape, dt=dt, spacing=g.spacing(), eps_r=eps_r, sigma=sigma,
                     mu_r=mu_r, cpml_n=layers, cpml_axes="xyz" if layers > 0 else "",
                     cpml_profiles=profiles, cpml_face_d=face_d, pec_axes="xyz",
                     pec_mask=pec, sources=sources, probes=probes, dft_planes=planes,
                     ports_vi=ports_vi, record_order="graded")
    cfg.flux, cfg.flux_area = _flux_entries(spec, g, n_steps, True)
    cfg.ntff = _ntff_entries(spec, g)
    if spec.get("subpixel") and not spec.get("debye") and not spec.get("lorentz"):
        # rfx/runners/nonuniform.py:648-672 (not with dispersion)
        cfg.aniso_eps = smoothed_eps_graded(g, _smoothing_boxes(spec))
    if spec.get("debye"):
        cfg.debye_poles, cfg.debye_masks = debye_pole_masks(
            shape, coords, spec.get("geometry", []), spec["debye"])
    if spec.get("lorentz"):
        cfg.lorentz_poles, cfg.lorentz_masks = debye_pole_masks(
            shape, coords, spec.get("geometry", []), spec["lorentz"])
    return g, cfg


def _ntff_entries(spec, g):
    """NTFF box of a spec: (corner_lo, corner_hi, freqs|None, n_freqs) ->
    [(lo, hi, freqs f32)] through the lane's position-to-index rule
    (rfx/farfield.py:127-141, rfx/runners/nonuniform.py:1086-1097); default
    frequencies as rfx/api/__init__.py:2666-2667, stored in float32."""
    ent = spec.get("ntff")
    if ent is None:
        return []
    corner_lo, corner_hi, freqs, nf = ent
    if freqs is None:
        freqs = np.linspace(spec["freq_max"] / 10, spec["freq_max"], nf)
    return [(g.position_to_index(corner_lo), g.position_to_index(corner_hi),
             np.asarray(freqs, F32))]


def _flux_entries(spec, g, n_steps, graded):
    """Flux monitors of a spec: (axis, coord, freqs|None, n_freqs, size|None,
    center|None, window, alpha) -> StepConfig.flux entries and dA arrays
    (rfx/runners/uniform.py:66-137; graded rfx/runners/nonuniform.py:924-958,
    full plane only, per-cell tangential widths)."""
    out, areas = [], []
    for ent in spec.get("flux_monitors", []):
        axis_name, coord, freqs, nf, size, center, window, alpha = ent
        a = "xyz".index(axis_name)
        pos = [0.0, 0.0, 0.0]
        pos[a] = coord
        index = g.position_to_index(tuple(pos))[a]
        if freqs is None:
            freqs = np.linspace(spec["freq_max"] / 10, spec["freq_max"], nf).astype(F32)
        tang = [t for t in range(3) if t != a]
        if size is None:
            (lo1, hi1), (lo2, hi2) = (0, g.shape[tang[0]]), (0, g.shape[tang[1]])
        else:
            if graded:
                raise NotImplementedError("finite flux regions on the graded lane")
            b = []
            dom = spec["domain"]
            for q, t in enumerate(tang):
                c = center[q] if center is not None else dom[t] / 2.0
                lo = max(0, int(round(c / g.dx - size[q] / (2.0 * g.dx))) + g.pad)
                hi = min(g.shape[t], int(round(c / g.dx + size[q] / (2.0 * g.dx))) + g.pad)
                b.append((lo, hi))
            (lo1, hi1), (lo2, hi2) = b
        if graded:
            w1, w2 = g.d32[tang[0]][lo1:hi1], g.d32[tang[1]][lo2:hi2]
        else:
            w1 = np.full(hi1 - lo1, F32(g.dx))
            w2 = np.full(hi2 - lo2, F32(g.dx))
        areas.append(w1.reshape(-1, 1) * w2.reshape(1, -1))
        out.append((a, index, lo1, hi1, lo2, hi2, np.asarray(freqs, F32),
                    dft_window_weights(n_steps, window, alpha)))
    return out, areas
