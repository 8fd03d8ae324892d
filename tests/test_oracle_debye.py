"""Pins of the oracle's Debye ADE (oracle/fdtd_np.py debye_pole_masks,
debye_coefficients, Stepper.update_e):

* the coefficients equal the hand-computed values of the reference's own
  test (tests/test_debye.py:38-80 there: dt 1 ps, tau 10 ps, d_eps 3,
  eps_inf 2);
* two exact limits of the Crank-Nicolson recursion: tau << dt makes
  P = eps0 d_eps E at every step, i.e. a plain dielectric of permittivity
  eps_inf + d_eps; tau >> dt leaves P at zero, i.e. eps_inf.  A wrong sign
  of cc, a P update from E^n alone or a missing term in gamma fails them;
* the pole masks follow the reference's assembly rules: OR-ed over shapes
  (a later box of another material does not remove the pole), equal poles
  merged, CPML pads not extended.
"""
import numpy as np

from oracle import fdtd_np as o
from tests import cases


def test_coefficients_match_the_reference_hand_values():
    dt, tau, de, einf = 1e-12, 10e-12, 3.0, 2.0
    shape = (2, 2, 2)
    ca, cb, alpha, beta, cc = o.debye_coefficients(
        [(de, tau)], [np.ones(shape, bool)], np.full(shape, einf, np.float32),
        np.zeros(shape, np.float32), dt)
    a = (2 * tau - dt) / (2 * tau + dt)
    b = o.EPS_0 * de * dt / (2 * tau + dt)
    g = o.EPS_0 * einf + b
    np.testing.assert_allclose(alpha[0], a, rtol=1e-12)
    np.testing.assert_allclose(beta[0], b, rtol=1e-12)
    np.testing.assert_allclose(ca, (o.EPS_0 * einf - b) / g, rtol=1e-12)
    np.testing.assert_allclose(cb, dt / g, rtol=1e-12)
    np.testing.assert_allclose(cc[0], (1 - a) / g, rtol=1e-12)


def _series(spec, n):
    g, cfg = o.build_uniform(spec, n)
    return o.run_loop(cfg, n).time_series


def test_fast_relaxation_is_the_static_dielectric():
    """tau = 1e-6 dt: the Debye box behaves as eps_inf + d_eps = 5."""
    n = 200
    dt = o.UniformGrid(10e9, (0.016, 0.012, 0.010), 1e-3, 4).dt
    deb = cases.debye_case(1e-6 * dt, overlap=False)
    plain = cases.debye_case(1e-6 * dt, overlap=False)
    plain.pop("debye")
    plain["materials"]["deb"] = (5.0, 0.01, 1.0)
    a, b = _series(deb, n), _series(plain, n)
    err = cases.rel_l2(a, b)
    print("fast-relaxation limit", err)
    assert err < 1e-4
    # a medium left at eps_inf differs by far
    plain["materials"]["deb"] = (2.0, 0.01, 1.0)
    assert cases.rel_l2(a, _series(plain, n)) > 1e-2


def test_slow_relaxation_is_eps_inf():
    """tau = 1e6 dt: the polarisation never builds up; the box is eps_inf."""
    n = 200
    dt = o.UniformGrid(10e9, (0.016, 0.012, 0.010), 1e-3, 4).dt
    deb = cases.debye_case(1e6 * dt, overlap=False)
    plain = cases.debye_case(1e6 * dt, overlap=False)
    plain.pop("debye")
    err = cases.rel_l2(_series(deb, n), _series(plain, n))
    print("slow-relaxation limit", err)
    assert err < 1e-4


def test_pole_masks_follow_the_assembly_rules():
    spec = cases.debye_case(1e-11)
    spec["materials"]["deb2"] = (4.0, 0.0, 1.0)
    # a second material with the same pole and one with another pole, the
    # latter touching the domain edge (the pad stays pole-free)
    spec["debye"]["deb2"] = [(3.0, 1e-11), (1.0, 5e-11)]
    spec["geometry"].append(((0.0, 0.0, 0.0), (0.003, 0.003, 0.003), "deb2"))
    g, cfg = o.build_uniform(spec, 1)
    assert cfg.debye_poles == [(3.0, 1e-11), (1.0, 5e-11)]
    m0, m1 = cfg.debye_masks
    coords = g.node_coords_f32()
    a = o.box_mask(coords, (0.004, 0.003, 0.002), (0.011, 0.009, 0.007))
    b = o.box_mask(coords, (0.009, 0.002, 0.001), (0.013, 0.006, 0.005))
    c = o.box_mask(coords, (0.0, 0.0, 0.0), (0.003, 0.003, 0.003))
    np.testing.assert_array_equal(m0, a | c)
    np.testing.assert_array_equal(m1, c)
    assert (a & b).any() and m0[a & b].all()          # the later box keeps the pole
    p = g.pad
    assert not m1[:p].any() and not m1[:, :p].any()  # pads not extended
    assert cfg.eps_r[p, p, p] == np.float32(4.0) and cfg.eps_r[0, 0, 0] == np.float32(4.0)
