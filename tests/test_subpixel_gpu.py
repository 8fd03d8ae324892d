"""Subpixel smoothing on an MI355X: runs with per-component (Kottke) eps at
the interface cells against the float64 oracle with the same smoothing,
uniform and graded; the smoothing visibly changes the result; the kernel
variants that do not carry per-component eps refuse such plans."""
import sys

import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-4


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


def _oracle(spec, n, graded):
    g, cfg = (o.build_graded if graded else o.build_uniform)(spec, n)
    assert cfg.aniso_eps is not None
    return cstep.run_c(cfg, n, want_fields=False)


@pytest.mark.parametrize("graded", [False, True])
def test_subpixel_run_vs_oracle(R, graded):
    spec = cases.mixed_graded_case() if graded else cases.mixed_case("run")
    n = 300
    _say(f"subpixel smoothing, {'graded' if graded else 'uniform'} mixed case, {n} steps")
    sim = cases.to_sim(spec)
    kw = {}
    if graded:
        kw["s_param_freqs"] = np.asarray(spec["s11_freqs"], np.float32)
    smooth = sim.run(n_steps=n, subpixel_smoothing=True, **kw)
    stair = sim.run(n_steps=n, **kw)
    spec = dict(spec, subpixel=True)
    ref = _oracle(spec, n, graded)
    e = cases.rel_l2(smooth.time_series, ref.time_series)
    d = cases.rel_l2(stair.time_series, ref.time_series)
    print("smoothed vs oracle", e, "staircase vs smoothed oracle", d)
    assert e <= TOL
    assert d > 10 * TOL


def test_subpixel_plans_refuse_other_variants(R):
    from rfx_amd import _lib
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    _say("subpixel smoothing: variant refusals and the autotuner")
    spec = cases.mixed_case("run")
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(40, "run", subpixel_smoothing=True)
    assert prob.mat.aniso
    dr = DeviceRun(prob, 40, "cuda")
    for v in ((1, 1), (2, 0), (7, 0), (8, 0)):
        with pytest.raises(_lib.RfxHipError):
            at.set_variant(dr, *v)
    res = at.autotune(dr, trial_steps=8)
    assert res["variant"][0] in (0, 4, 5, 6)
    dr.reset()
    dr.run(0, 40)
    series = {}
    for v in ((0, 1), (4, 0), (6, 1)):
        at.set_variant(dr, *v)
        dr.reset()
        dr.run(0, 40)
        series[v] = dr.time_series().cpu().numpy()
    assert np.array_equal(series[(0, 1)], series[(4, 0)])
    assert np.array_equal(series[(0, 1)], series[(6, 1)])
    dr.close()
