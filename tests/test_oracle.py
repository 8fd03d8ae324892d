"""Pins the CPU oracle: golden vectors, closed forms, C == numpy.

These tests exercise only the oracle (test infrastructure); the HIP path
is compared with the oracle in test_parity_gpu.py.
"""
import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from tests import cases


def test_golden_s11_numpy_and_c():
    """rfx golden S11 (tests/test_lumped_rlc_ad.py:135-151), rtol 1e-5."""
    golden = cases.golden_s11()
    g, cfg = o.build_uniform(cases.S11_CASE, cases.S11_STEPS)
    assert g.shape == (28, 28, 28)
    for out in (cstep.run_c(cfg, cases.S11_STEPS, want_fields=False),):
        s11 = o.extract_lumped_s11(out.port_v_dft[0], out.port_i_dft[0])
        np.testing.assert_allclose(s11, golden, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("rung", [0, 1])
def test_ladder_objective_values(rung):
    """J0, J+, J- of the gradient ladder fixture to 1e-5 relative."""
    lad = cases.ladder()["rungs"][rung]
    n = lad["lambda_over_dx"]
    h = lad["fd_step_abs"]
    vals = {}
    for key, e in (("J0", 4.0), ("J_plus", 4.0 + h), ("J_minus", 4.0 - h)):
        spec, n_steps, _ = cases.ladder_case(n, e)
        assert n_steps == lad["n_steps"]
        g, cfg = o.build_uniform(spec, n_steps)
        assert list(g.shape) == lad["grid_shape"]
        out = cstep.run_c(cfg, n_steps, want_fields=False)
        vals[key] = cases.ladder_objective(out.time_series[:, 0], g.dt)
        assert abs(vals[key] - lad[key]) / lad[key] < 1e-5, (key, vals[key], lad[key])
    fd = (vals["J_plus"] - vals["J_minus"]) / (2 * h)
    # the float64 oracle's central difference agrees with the reference's
    # float32 jax.grad to the fixture's own AD-vs-FD level
    assert abs(fd - lad["grad_ad"]) / abs(lad["grad_ad"]) < 2e-4


def test_c_equals_numpy_mixed_case():
    g, cfg = o.build_uniform(cases.mixed_case("run"), 120)
    a = o.run_loop(cfg, 120)
    b = cstep.run_c(cfg, 120)
    np.testing.assert_array_equal(a.time_series, b.time_series)
    for q in range(len(a.dft_planes)):
        np.testing.assert_array_equal(a.dft_planes[q], b.dft_planes[q])
    for c in o.COMPS:
        np.testing.assert_array_equal(a.fields[c], b.fields[c])
    np.testing.assert_allclose(a.port_v_dft[0], b.port_v_dft[0], rtol=1e-12)


def test_dft_plane_equals_direct_dft_of_probe():
    """A DFT-plane bin equals the direct DFT of the probe series at the same
    cell with t = (n+1) dt (rfx/simulation.py:1574-1588, test_dft_probes.py)."""
    spec = cases.mixed_case("run")
    spec["probes"] = [((0.008, 0.006, 0.005), "ez")]
    freqs = np.array([3e9, 6e9, 9e9], np.float32)
    spec["dft_planes"] = [("z", 0.005, "ez", freqs, 3)]
    n = 150
    g, cfg = o.build_uniform(spec, n)
    out = cstep.run_c(cfg, n, want_fields=False)
    (i, j, k), _ = cfg.probes[0]
    ts = out.time_series[:, 0]
    t = (np.arange(n) + 1) * g.dt
    for f_i, f in enumerate(freqs.astype(np.float64)):
        direct = np.sum(ts * np.exp(-1j * 2 * np.pi * f * t)) * g.dt
        np.testing.assert_allclose(out.dft_planes[0][f_i, i, j], direct, rtol=1e-12)


def test_zero_source_stays_zero():
    spec = cases.mixed_case("run")
    spec["ports"] = []
    g, cfg = o.build_uniform(spec, 50)
    out = cstep.run_c(cfg, 50)
    for c in o.COMPS:
        assert not out.fields[c].any()


def test_cpml_profile_closed_form():
    """rfx/boundaries/cpml.py:101-171: sigma_max, b = exp(-(s/k+a)dt/eps0),
    outermost cell alpha = 0 so c = b - 1 when kappa = 1; innermost sigma = 0
    so c = 0 and b = exp(-alpha dt/eps0) with alpha = 0.05."""
    n, dt, d = 8, 1e-12, 1e-3
    b, c, k = o.cpml_profile(n, dt, d)
    eta = np.sqrt(o.MU_0 / o.EPS_0)
    smax = -np.log(1e-15) * 4 / (2 * eta * n * d)
    assert np.isclose(b[0], np.exp(-smax * dt / o.EPS_0), rtol=1e-6)
    assert np.isclose(c[0], b[0] - 1.0, rtol=1e-6)
    assert c[-1] == 0.0 and np.isclose(b[-1], np.exp(-0.05 * dt / o.EPS_0), rtol=1e-6)
    assert np.all(k == 1.0)
    # sigma grading rho^3: c/(b-1) = sigma/(sigma + alpha) recovers sigma
    rho = 1 - np.arange(n) / (n - 1)
    sig = smax * rho ** 3
    alpha = 0.05 * (1 - rho)
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = np.where(sig > 0, sig / (sig + alpha), 0.0)
    np.testing.assert_allclose(np.where(sig > 0, c / (b - 1.0), 0.0), ratio, rtol=1e-5)


def test_box_rasterisation_brute_force():
    """Per-node check of the half-open [lo, hi) float32 rule
    (rfx/geometry/csg.py:243-302) and the thin-sheet branch."""
    g = o.UniformGrid(5e9, (0.14, 0.06, 0.05), 2e-3, 8)
    coords = g.node_coords_f32()
    lo, hi = (0.07, 0.018, 0.018), (0.09, 0.042, 0.032)
    m = o.box_mask(coords, lo, hi)
    for a in range(3):
        expect = [(np.float32(c) >= np.float32(lo[a])) and (np.float32(c) < np.float32(hi[a]))
                  for c in coords[a]]
        axis_any = m.any(axis=tuple(b for b in range(3) if b != a))
        assert list(axis_any) == expect
    # README box: x nodes 43..52, y 17..28, z 17..23 (incl. 8-cell pads)
    idx = np.argwhere(m)
    assert tuple(idx.min(0)) == (43, 17, 17) and tuple(idx.max(0)) == (52, 28, 23)
    # a sub-cell thin sheet snaps to the single nearest node
    thin = o.box_mask(coords, (0.01, 0.01, 0.0203), (0.02, 0.02, 0.0207))
    zs = np.unique(np.argwhere(thin)[:, 2])
    assert list(zs) == [int(np.argmin(np.abs(coords[2] - np.float32(0.0205))))]


def test_pad_extension_hi_face_fallback():
    """rfx/geometry/rasterize_grid.py:237-414: a box touching the hi face
    loses its last node to the half-open rule; the pad and that node are
    refilled from one slice inward."""
    shape = (12, 5, 5)
    e = np.ones(shape, np.float32)
    e[2:9] = 4.0  # interior spans nodes 2..9; node 9 is the last interior node
    s = np.zeros(shape, np.float32)
    m = np.ones(shape, np.float32)
    e2, _, _ = o.extend_cpml_pad(e, s, m, 2)
    assert np.all(e2[:, 2:3] == e2[:, 1:2])
    assert np.all(e2[9:12, 2, 2] == 4.0)       # node 9 repaired, pads 10,11 filled
    assert np.all(e2[0:2, 2, 2] == 4.0)        # lo pad copies node 2
