"""CFS-CPML profiles (rfx/boundaries/cpml.py:101-171, 346-478).

Per face the kernels need, in slab-local order (index 0 = the slab cell
with the lowest coordinate): b, c and 1/kappa - 1, plus the inverse face
cell size.  Lo faces use the profile as computed (outermost cell first);
hi faces use it reversed, which is the reference's pre-flipped hi profile
(cpml.py:404-409).  Profiles are computed in float64 and rounded to
float32, as the reference does.
"""
from __future__ import annotations

import numpy as np

from .constants import EPS_0, MU_0


def cpml_profile(n_layers: int, dt: float, d: float, kappa_max: float = 1.0,
                 order: int = 3, R: float = 1e-15):
    eta = float(np.sqrt(MU_0 / EPS_0))
    sigma_max = -float(np.log(R)) * (order + 1) / (2.0 * eta * n_layers * d) * kappa_max
    rho = 1.0 - np.arange(n_layers, dtype=np.float64) / max(n_layers - 1, 1)
    sigma = sigma_max * rho ** order
    kappa = 1.0 + (kappa_max - 1.0) * rho ** order
    alpha = 0.05 * (1.0 - rho)
    denom = sigma * kappa + kappa ** 2 * alpha
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS_0)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(denom > 1e-30, sigma * (b - 1.0) / denom, 0.0)
    return b.astype(np.float32), c.astype(np.float32), kappa.astype(np.float32)


def face_tables(n_layers: int, dt: float, face_sizes, kappa_max: float = 1.0,
                axes: str = "xyz"):
    """Six faces (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi) -> (n, b, c, km1, inv_d)."""
    faces = []
    for ax in range(3):
        for side in range(2):
            if n_layers <= 0 or "xyz"[ax] not in axes:
                faces.append((0, None, None, None, 0.0))
                continue
            d = float(face_sizes[ax][side])
            b, c, k = cpml_profile(n_layers, dt, d, kappa_max)
            if side == 1:
                b, c, k = b[::-1].copy(), c[::-1].copy(), k[::-1].copy()
            km1 = (np.float32(1.0) / k - np.float32(1.0)).astype(np.float32)
            faces.append((n_layers, b, c, km1, float(np.float32(1.0) / np.float32(d))))
    return faces
