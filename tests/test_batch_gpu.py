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
    from rfx_amd.batch import run_batch
    rng = np.random.default_rng(5)
    shapes = [(24, 24, 24), (32, 20, 16), (16, 28, 30), (40, 16, 12), (20, 20, 36)]
    specs = [cases.sweep_instance(rng, s) for s in shapes]
    specs.append(cases.mixed_graded_case())
    steps = [300, 120, 251, 37, 300, 190]
    sims = [cases.to_sim(sp) for sp in specs]
    res = run_batch(sims, steps, use_graph=use_graph)
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
