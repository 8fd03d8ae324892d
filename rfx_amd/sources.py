"""Source waveforms and port helpers (rfx/sources/sources.py).

Waveform tables are evaluated in float32 on float32 sample times
f32(n) * f32(dt), as the reference's jax.vmap over
jnp.arange(n_steps, dtype=float32) * dt does (rfx/simulation.py:273).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

F32 = np.float32


@dataclass(frozen=True)
class GaussianPulse:
    """Differentiated Gaussian s(t) = A (-2u) exp(-u^2), u = (t - t0)/tau,
    tau = 1/(f0 bw pi), t0 = cutoff tau (rfx/sources/sources.py:56-104)."""

    f0: float
    bandwidth: float = 0.5
    amplitude: float = 1.0
    cutoff: float = 3.0

    @property
    def tau(self) -> float:
        return 1.0 / (self.f0 * self.bandwidth * math.pi)

    @property
    def t0(self) -> float:
        return self.cutoff * self.tau

    def __call__(self, t):
        t = np.asarray(t, dtype=F32)
        arg = (t - F32(self.t0)) / F32(self.tau)
        return (F32(self.amplitude) * (F32(-2.0) * arg) * np.exp(-(arg ** 2))).astype(F32)


def sample_times(n_steps: int, dt: float) -> np.ndarray:
    return np.arange(n_steps, dtype=F32) * F32(dt)


@dataclass(frozen=True)
class LumpedPort:
    position: tuple
    component: str
    impedance: float
    excitation: GaussianPulse


def port_sigma(cell_sizes, component: str, impedance: float) -> float:
    """sigma = d_par / (Z0 d_perp1 d_perp2) (rfx/sources/sources.py:31-45)."""
    a = {"ex": 0, "ey": 1, "ez": 2}[component]
    perp = [cell_sizes[i] for i in range(3) if i != a]
    return cell_sizes[a] / (impedance * perp[0] * perp[1])
