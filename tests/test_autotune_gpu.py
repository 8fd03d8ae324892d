"""Kernel variants and the autotuner on an MI355X: every variant (two-pass
or fused step, flat or XCD-band workgroup order, marching step with any
segment count) meets the oracle bar on the mixed case and the graded case,
and the tuner leaves a valid choice and a zeroed state."""
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


def test_every_variant_matches_oracle_and_tuner_picks_one(R):
    import torch
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    _say("variants x mixed case (run lane), autotune")
    spec = cases.mixed_case("run")
    n = 300
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(n, "run")
    g, cfg = o.build_uniform(spec, n)
    ref = cstep.run_c(cfg, n, want_fields=False)
    dr = DeviceRun(prob, n, "cuda")
    series = {}
    for v in at.variants(dr):
        at.set_variant(dr, *v)
        assert at.get_variant(dr) == v
        dr.reset()
        dr.run(0, n)
        torch.cuda.synchronize()
        ts = dr.time_series().cpu().numpy()
        e = cases.rel_l2(ts, ref.time_series)
        print("variant", v, e)
        assert e <= 1e-4
        series[v] = ts
    base = series[(0, 1)]
    assert np.array_equal(series[(0, 0)], base)          # order does not change arithmetic
    for v in series:
        if v[0] in (4, 5, 6):                             # columns == one cell per thread
            assert np.array_equal(series[v], base), v
        if v[0] in (2, 3, 7, 8):                          # marching == fused arithmetic per cell
            assert cases.rel_l2(series[v], series[(1, 1)]) <= 1e-6
    res = at.autotune(dr, trial_steps=16)
    print("autotune", res)
    assert tuple(res["variant"]) in at.variants(dr)
    assert not any(float(f.abs().max()) for f in dr.fields)
    dr.run(0, n)
    torch.cuda.synchronize()
    assert cases.rel_l2(dr.time_series().cpu().numpy(), ref.time_series) <= 1e-4
    dr.close()


def test_march_graded_vs_oracle(R):
    import torch
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    _say("marching kernel on the graded mixed case")
    spec = cases.mixed_graded_case()
    n = 300
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(n, "run", s11_freqs=np.asarray(spec["s11_freqs"], np.float32))
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, n, want_fields=False)
    dr = DeviceRun(prob, n, "cuda")
    for kern, seg in ((2, 0), (2, 1), (2, 3), (3, 0), (3, 7), (4, 1), (5, 0), (6, 1), (7, 0),
                      (7, 1), (7, 5), (8, 0), (8, 2)):
        at.set_variant(dr, kern, seg)
        dr.reset()
        dr.run(0, n)
        torch.cuda.synchronize()
        ts = dr.time_series().cpu().numpy()
        for p in range(ts.shape[1]):
            e = cases.rel_l2(ts[:, p], ref.time_series[:, p])
            print("march", kern, "seg", seg, "probe", p, e)
            assert e <= 1e-4
        pv = dr.port_samples().cpu().numpy()[:, 0, 0]
        assert cases.rel_l2(pv, ref.port_v[0]) <= 1e-4
    dr.close()


def test_batch_autotune(R):
    from rfx_amd import autotune as at
    from rfx_amd.batch import Batch
    _say("batch autotune")
    rng = np.random.default_rng(9)
    specs = [cases.sweep_instance(rng, s) for s in [(24, 24, 24), (32, 20, 16), (16, 28, 30)]]
    sims = [cases.to_sim(sp) for sp in specs]
    b = Batch(sims, [120, 80, 100])
    res = at.autotune_batch(b, trial_steps=16)
    print("batch autotune", res)
    out = b.run()
    for q, (sim, n) in enumerate(zip(sims, [120, 80, 100])):
        ind = sim.run(n_steps=n)
        assert cases.rel_l2(out[q].time_series, ind.time_series) <= 1e-6
    b.close()
