"""Pins of the oracle's Lorentz ADE (oracle/fdtd_np.py lorentz_coefficients,
Stepper.update_e with Lorentz and mixed poles):

* the coefficients equal the hand-computed values of the reference's own
  test (tests/test_lorentz.py:47-75 there);
* the exact discrete dispersion relation: a lossless PEC cavity filled with
  a lossless Lorentz medium rings at the two polariton frequencies of its
  TM110 mode, Omega^2 (eps_inf + kappa/(w0^2 - Omega^2)) / c^2 = K^2 with
  Omega = (2/dt) sin(w dt/2) and K the Yee wavenumber of the mode (the
  central-difference recursion gives chi = kappa / (w0^2 - Omega^2)
  exactly); a wrong sign of the coupling, P^{n+1} taken from E^{n+1} or a
  wrong c fails it;
* a Lorentz pole with kappa = 0 leaves a plain eps_inf medium, and next to
  Debye poles it leaves the Debye run: this pins the mixed coefficients
  (gamma_t = 1/cc_L + sum beta) of rfx/simulation.py:470-478.
"""
import numpy as np

from oracle import fdtd_np as o
from tests import cases
from tests.spectral import matrix_pencil_freqs


def test_coefficients_match_the_reference_hand_values():
    dt, w0, delta = 1e-12, 2 * np.pi * 10e9, 1e9
    kappa = 3.0 * w0 ** 2
    shape = (2, 2, 2)
    ca, cb, a, b, c, cc = o.lorentz_coefficients(
        [(w0, delta, kappa)], [np.ones(shape, bool)], np.ones(shape, np.float32),
        np.zeros(shape, np.float32), dt)
    denom = 1.0 + delta * dt
    np.testing.assert_allclose(a[0], (2.0 - w0 ** 2 * dt ** 2) / denom, rtol=1e-12)
    np.testing.assert_allclose(b[0], -(1.0 - delta * dt) / denom, rtol=1e-12)
    np.testing.assert_allclose(c[0], o.EPS_0 * kappa * dt ** 2 / denom, rtol=1e-12)
    np.testing.assert_allclose(cc, 1.0 / o.EPS_0, rtol=1e-12)
    np.testing.assert_allclose(cb, dt / o.EPS_0, rtol=1e-12)


def polariton_freqs(eps_inf, w0, kappa, a, b, dx, dt):
    """TM110 frequencies of a PEC cavity filled with a lossless Lorentz medium
    on the Yee grid (exact discrete dispersion relation)."""
    c2 = 1.0 / (o.EPS_0 * o.MU_0)
    K2 = (2 / dx * np.sin(np.pi * dx / (2 * a))) ** 2 + (2 / dx * np.sin(np.pi * dx / (2 * b))) ** 2
    B = eps_inf * w0 ** 2 + kappa + K2 * c2
    disc = np.sqrt(B * B - 4 * eps_inf * K2 * c2 * w0 ** 2)
    om2 = np.array([(B - disc) / (2 * eps_inf), (B + disc) / (2 * eps_inf)])
    return 2 / dt * np.arcsin(np.sqrt(om2) * dt / 2) / (2 * np.pi)


def test_lorentz_cavity_polaritons():
    spec, n, pred = cases.lorentz_cavity_case()
    g, cfg = o.build_uniform(spec, n)
    out = o.run_loop(cfg, n)
    f, _ = matrix_pencil_freqs(out.time_series[:, 0], cfg.dt, max_order=60)
    err = [np.min(np.abs(f - p)) / p for p in pred]
    print("predicted", pred, "found", [f[np.argmin(np.abs(f - p))] for p in pred], err)
    assert max(err) < 1e-5


def _series(spec, n):
    g, cfg = o.build_uniform(spec, n)
    return o.run_loop(cfg, n).time_series


def test_zero_coupling_leaves_the_plain_and_the_debye_runs():
    n = 150
    w0 = 2 * np.pi * 9e9
    spec = cases.debye_case(20e-12, overlap=False)
    spec["lorentz"] = {"deb": [(w0, 1e9, 0.0)]}
    debye_only = cases.debye_case(20e-12, overlap=False)
    e = cases.rel_l2(_series(spec, n), _series(debye_only, n))
    print("Debye + zero Lorentz vs Debye", e)
    assert e < 1e-12
    lor = cases.debye_case(20e-12, overlap=False)
    lor.pop("debye")
    lor["lorentz"] = {"deb": [(w0, 1e9, 0.0)]}
    plain = cases.debye_case(20e-12, overlap=False)
    plain.pop("debye")
    e = cases.rel_l2(_series(lor, n), _series(plain, n))
    print("zero Lorentz vs plain", e)
    assert e < 1e-12
