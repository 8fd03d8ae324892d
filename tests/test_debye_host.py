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
    from rfx_amd.materials import POLE_SHIFT, debye_tables
    spec = _spec()
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(10, "run")
    g, cfg = o.build_uniform(spec, 10)
    assert [tuple(p) for p in prob.poles] == cfg.debye_poles
    coef, flags, rows = debye_tables(prob.mat, prob.dt, prob.poles)
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


def test_lorentz_and_gradients_fail_loudly():
    import pytest
    import rfx_amd as R
    sim = R.Simulation(10e9, (0.01, 0.01, 0.01), dx=1e-3, cpml_layers=4)
    with pytest.raises(NotImplementedError):
        sim.add_material("l", eps_r=2.0, lorentz_poles=[(1.0, 1e9, 1e8)])
    with pytest.raises(ValueError):
        sim.add_material("bad", eps_r=2.0, debye_poles=[(1.0, 0.0)])
