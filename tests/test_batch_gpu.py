"""Batched sweep of uneven instances on an MI355X: every instance of the
batch equals its own independent run (same kernels, same arithmetic), and
a uniform instance matches the float64 oracle."""
import sys

import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from tests import cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def R():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rfx_amd
    return rfx_amd


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_independent_runs(R, use_graph):
    _say(f"batch of 6 uneven instances (graph={use_graph})")
    from rfx_amd.batch import run_sims
    rng = np.random.default_rng(5)
    shapes = [(24, 24, 24), (32, 20, 16), (16, 28, 30), (40, 16, 12), (20, 20, 36)]
    specs = [cases.sweep_instance(rng, s) for s in shapes]
    specs.append(cases.mixed_graded_case())
    steps = [300, 120, 251, 37, 300, 190]
    sims = [cases.to_sim(sp) for sp in specs]
    res = run_sims(sims, steps, use_graph=use_graph)
    for q, (sim, n) in enumerate(zip(sims, steps)):
        ind = sim.run(n_steps=n)
        a, b = res[q].time_series, ind.time_series
        assert a.shape == (n, b.shape[1])
        exact = np.array_equal(a, b)
        e = cases.rel_l2(a, b)
        print("instance", q, shapes[q] if q < len(shapes) else "graded", n, "bit-exact", exact, e)
        assert e <= 1e-6
        for name in ind.dft_planes:
            np.testing.assert_allclose(res[q].dft_planes[name].accumulator,
                                       ind.dft_planes[name].accumulator, rtol=1e-6, atol=1e-30)
    g, cfg = o.build_uniform(specs[2], steps[2])
    ref = cstep.run_c(cfg, steps[2], want_fields=False)
    e = cases.rel_l2(res[2].time_series, ref.time_series)
    print("instance 2 vs oracle", e)
    assert e <= 1e-4


def test_batch_mixed_layouts(R):
    """Instances on one cell per thread and on 2-cell columns, in flat and
    XCD-band workgroup order, share one batched launch; each equals its own
    run."""
    _say("batch with mixed per-instance layouts")
    from rfx_amd import autotune as at
    from rfx_amd.batch import Batch
    rng = np.random.default_rng(11)
    shapes = [(24, 24, 24), (33, 20, 16), (15, 28, 30), (40, 16, 12)]
    specs = [cases.sweep_instance(rng, s) for s in shapes]
    specs.append(cases.mixed_graded_case())
    steps = [200, 120, 151, 77, 160]
    layouts = [(0, 1), (4, 0), (4, 1), (0, 0), (4, 1)]
    sims = [cases.to_sim(sp) for sp in specs]
    b = Batch(sims, steps)
    try:
        for r, v in zip(b.runs, layouts):
            at.set_variant(r, *v)
        res = b.run()
    finally:
        b.close()
    for q, (sim, n) in enumerate(zip(sims, steps)):
        ind = sim.run(n_steps=n)
        e = cases.rel_l2(res[q].time_series, ind.time_series)
        print("instance", q, at.variant_name(layouts[q]), n, e)
        assert e <= 1e-6


def _slab_spec(eps=3.0, f0=8e9, sig=0.01):
    """A small CPML box with a dielectric slab, a second material box painted
    over part of it, an Ez source, two probes and a DFT plane."""
    return dict(freq_max=20e9, domain=(0.024, 0.020, 0.018), dx=1e-3, boundary="cpml",
                cpml_layers=6,
                materials={"slab": (eps, sig, 1.0), "cap": (2.0, 0.0, 1.0)},
                geometry=[((0.010, 0.0, 0.0), (0.016, 0.020, 0.018), "slab"),
                          ((0.012, 0.006, 0.005), (0.014, 0.012, 0.011), "cap")],
                ports=[((0.005, 0.010, 0.009), "ez", 0.0, o.Pulse(f0, 0.7), True)],
                probes=[((0.019, 0.010, 0.009), "ez"), ((0.008, 0.012, 0.006), "ex")],
                dft_planes=[("x", 0.018, "ez", None, 3)], lane="run")


def test_run_batch_reference_signature(R):
    """run_batch(sim_factory, sweep, run_kwargs=...) -> [(params, Result)] in
    sweep order, each equal to the factory's own sim.run(); batched when
    run_kwargs holds only n_steps, sequential otherwise."""
    from rfx_amd.batch import ParameterSweep, run_batch
    _say("run_batch with a parameter sweep")
    sweep = ParameterSweep(eps=[2.0, 4.5], f0=[6e9, 9e9])

    def factory(eps, f0):
        return cases.to_sim(_slab_spec(eps, f0))

    n = 160
    out = run_batch(factory, sweep, run_kwargs={"n_steps": n})
    seq = run_batch(factory, sweep, run_kwargs={"n_steps": n, "autotune": False})
    assert [p for p, _ in out] == list(sweep.combinations()) == [p for p, _ in seq]
    for (params, res), (_, res2) in zip(out, seq):
        ind = factory(**params).run(n_steps=n)
        assert cases.rel_l2(res.time_series, ind.time_series) <= 1e-6
        assert cases.rel_l2(res2.time_series, ind.time_series) <= 1e-6
        for name in ind.dft_planes:
            np.testing.assert_allclose(res.dft_planes[name].accumulator,
                                       ind.dft_planes[name].accumulator, rtol=1e-6, atol=1e-30)
    params, res = out[3]
    g, cfg = o.build_uniform(_slab_spec(params["eps"], params["f0"]), n)
    ref = cstep.run_c(cfg, n, want_fields=False)
    assert cases.rel_l2(res.time_series, ref.time_series) <= 1e-4


def test_vmap_material_sweep(R):
    """vmap_material_sweep: a named material's eps_r (inside its shapes, the
    later box painted over it included, as rfx/vmap_sweep.py:314-360 does),
    each instance vs the float64 oracle with the same sweep rule; a global
    sigma sweep vs the simulation with that sigma; output shapes."""
    from oracle import cstep as cs
    from rfx_amd.batch import vmap_material_sweep
    _say("vmap_material_sweep slab.eps_r and sigma")
    n = 150
    sim = cases.to_sim(_slab_spec())
    vals = [2.0, 3.5, 6.0]
    r = vmap_material_sweep(sim, "slab.eps_r", vals, n_steps=n)
    assert r.time_series.shape == (3, n, 2)
    assert r.dft_planes is not None
    (name, acc), = r.dft_planes.items()
    assert acc.shape[0] == 3 and acc.shape[1] == 3
    for q, v in enumerate(vals):
        spec = _slab_spec()
        spec["material_sweep"] = ("slab", "eps_r", v)
        g, cfg = o.build_uniform(spec, n)
        ref = cs.run_c(cfg, n, want_fields=False)
        err = cases.rel_l2(r.time_series[q], ref.time_series)
        print("slab.eps_r", v, err)
        assert err <= 1e-4
    rs = vmap_material_sweep(sim, "sigma", [0.0, 0.05], n_steps=n)
    ind = cases.to_sim(_slab_spec(sig=0.05)).run(n_steps=n)
    assert cases.rel_l2(rs.time_series[1], ind.time_series) <= 1e-6
    with pytest.raises(ValueError):
        vmap_material_sweep(sim, "slab.eps_r", [2.0], n_steps=n, return_fields=True)
