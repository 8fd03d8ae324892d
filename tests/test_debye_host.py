"""Host set-up of Debye materials (CPU): the pole bits the Simulation folds
into the material ids equal the oracle's per-pole masks cell by cell, and
the per-material rows (alpha, beta, cc) and the Debye ca, cb equal the
oracle's coefficients (float32 against float64)."""
import numpy as np

from oracle import fdtd_np as o
from tests import cases


def _spec():
    spec = cases.debye_case(1e-11)
    spec["materials"]["deb2"] = (4.0, 0.0, 1.0)
    spec["debye"]["deb2"] = [(3.0, 1e-11), (1.0, 5e-11)]
    spec["geometry"].append(((0.0, 0.0, 0.0), (0.003, 0.003, 0.003), "deb2"))
    return spec


def test_pole_bits_and_rows_match_the_oracle():
    from rfx_amd.materials import POLE_SHIFT, dispersion_tables
    spec = _spec()
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(10, "run")
    g, cfg = o.build_uniform(spec, 10)
    assert [tuple(p) for p in prob.poles] == cfg.debye_poles
    coef, flags, rows = dispersion_tables(prob.mat, prob.dt, prob.poles, prob.lorentz)
    ids = prob.mat.ids
    for q, m in enumerate(cfg.debye_masks):
        got = ((flags[ids] >> (POLE_SHIFT + q)) & 1).astype(bool)
        np.testing.assert_array_equal(got, m)
    ca, cb, alpha, beta, cc = o.debye_coefficients(cfg.debye_poles, cfg.debye_masks,
                                                   cfg.eps_r, cfg.sigma, cfg.dt)
    np.testing.assert_allclose(coef[ids, 0], ca, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(coef[ids, 1], cb, rtol=2e-6)
    for q in range(len(cfg.debye_poles)):
        np.testing.assert_allclose(rows[ids, q, 0], alpha[q], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(rows[ids, q, 1], beta[q], rtol=1e-6, atol=1e-30)
        np.testing.assert_allclose(rows[ids, q, 2], cc[q], rtol=2e-6)


def test_bad_poles_fail_loudly():
    import pytest
    import rfx_amd as R
    sim = R.Simulation(10e9, (0.01, 0.01, 0.01), dx=1e-3, cpml_layers=4)
    with pytest.raises(NotImplementedError):
        sim.add_material("k", eps_r=2.0, chi3=1e-20)
    with pytest.raises(ValueError):
        sim.add_material("bad", eps_r=2.0, debye_poles=[(1.0, 0.0)])


def test_mixed_debye_lorentz_tables_match_the_oracle():
    """Debye and Lorentz poles together: the Lorentz rows (a, b, c, cc_L) and
    the mixed ca, cb, cc_D (rfx/simulation.py:470-478) against the oracle's
    stepper coefficients."""
    from rfx_amd.materials import dispersion_tables
    spec = _spec()
    w0 = 2 * np.pi * 9e9
    spec["lorentz"] = {"deb": [(w0, 1e9, 1.2 * w0 ** 2)], "d": [(0.0, 5e9, 4e20)]}
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(10, "run")
    g, cfg = o.build_uniform(spec, 10)
    assert [tuple(p) for p in prob.lorentz] == cfg.lorentz_poles
    coef, flags, rows = dispersion_tables(prob.mat, prob.dt, prob.poles, prob.lorentz)
    st = o.Stepper(cfg)
    ids = prob.mat.ids
    nd = len(prob.poles)
    np.testing.assert_allclose(coef[ids, 0], st.ca, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(coef[ids, 1], st.cb, rtol=2e-6)
    alpha, beta, cc, _ = st.debye
    for q in range(nd):
        np.testing.assert_allclose(rows[ids, q, 2], cc[q], rtol=2e-6)
    a_l, b_l, c_l, cc_l, _, _ = st.lorentz
    for q in range(len(prob.lorentz)):
        np.testing.assert_allclose(rows[ids, nd + q, 0], a_l[q], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(rows[ids, nd + q, 1], b_l[q], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(rows[ids, nd + q, 2], c_l[q], rtol=1e-6, atol=1e-40)
        np.testing.assert_allclose(rows[ids, nd + q, 3], cc_l, rtol=2e-6)
