"""Host logic of the reference's batch entry points (no GPU): parameter
sweeps (rfx/batch.py:39-72), parameter names (rfx/vmap_sweep.py:108-125)
and the material arrays of a material sweep, cell for cell against the
oracle's restatement of rfx/vmap_sweep.py:265-390."""
import numpy as np
import pytest

from oracle import fdtd_np as o
from tests import cases
from tests.test_batch_gpu import _slab_spec


def test_parameter_sweep_is_the_cartesian_product():
    from rfx_amd.batch import ParameterSweep
    sw = ParameterSweep(width=[0.01, 0.02], eps_r=np.array([2.0, 4.0, 6.0]))
    assert sw.keys == ["width", "eps_r"]
    assert sw.total == 6
    combos = list(sw.combinations())
    assert combos[0] == {"width": 0.01, "eps_r": 2.0}
    assert combos[1] == {"width": 0.01, "eps_r": 4.0}
    assert combos[-1] == {"width": 0.02, "eps_r": 6.0}
    assert all(type(v) is float for c in combos for v in c.values())


def test_param_names():
    from rfx_amd.batch import _parse_param_name
    assert _parse_param_name("eps_r") == (None, "eps_r")
    assert _parse_param_name("substrate.sigma") == ("substrate", "sigma")
    assert _parse_param_name("a.b.mu_r") == ("a.b", "mu_r")
    with pytest.raises(ValueError):
        _parse_param_name("substrate.thickness")


@pytest.mark.parametrize("sweep", [("slab", "eps_r", 5.5), ("cap", "sigma", 0.2),
                                   (None, "eps_r", 7.0), (None, "sigma", 0.03),
                                   (None, "mu_r", 1.5)])
def test_swept_materials_equal_the_oracle_rule(sweep):
    spec = _slab_spec()
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(10, "run", material_sweep=sweep)
    e, s, m = prob.mat.cell_arrays()
    spec["material_sweep"] = sweep
    g, cfg = o.build_uniform(spec, 10)
    np.testing.assert_array_equal(e, cfg.eps_r)
    np.testing.assert_array_equal(s, cfg.sigma)
    np.testing.assert_array_equal(m, cfg.mu_r)


def test_named_sweep_reaches_the_later_box_and_the_pads():
    """The oracle rule itself: a named sweep covers every cell of that
    material's shapes, the box painted over it later included, and the
    pads extended from them; the global rule leaves vacuum alone."""
    spec = _slab_spec()
    g, base = o.build_uniform(spec, 10)
    spec["material_sweep"] = ("slab", "eps_r", 5.5)
    _, cfg = o.build_uniform(spec, 10)
    cap = base.eps_r == np.float32(2.0)          # the later box
    assert cap.any() and np.all(cfg.eps_r[cap] == np.float32(5.5))
    slab = base.eps_r == np.float32(3.0)
    assert np.all(cfg.eps_r[slab] == np.float32(5.5))
    vac = base.eps_r == np.float32(1.0)
    assert np.all(cfg.eps_r[vac] == np.float32(1.0))
    p = 6
    assert np.all(cfg.eps_r[:, :p][slab[:, :p] | cap[:, :p]] == np.float32(5.5))
    spec["material_sweep"] = (None, "eps_r", 9.0)
    _, cfg = o.build_uniform(spec, 10)
    assert np.all(cfg.eps_r[~vac] == np.float32(9.0)) and np.all(cfg.eps_r[vac] == 1.0)


def test_batch_segment_steps():
    """Segment length of a batched reverse run: all steps when the record
    fits, else the largest S whose records and checkpoints fit, else
    MemoryError."""
    import math
    from rfx_amd.adjoint import batch_segment_steps
    T, per, st = [100, 70, 40], [10, 20, 30], [500, 400, 300]
    assert batch_segment_steps(T, per, st, 100 * 60) == 100

    def need(S):
        return S * sum(per) + sum((math.ceil(t / S) - 1) * b for t, b in zip(T, st))

    budget = 5000
    S = batch_segment_steps(T, per, st, budget)
    assert need(S) <= budget
    assert all(need(s) > budget for s in range(S + 1, 101))
    assert batch_segment_steps(T, per, st, budget, forced=7) == 7
    with pytest.raises(MemoryError):
        batch_segment_steps(T, per, st, 100)
