"""The temporal-blocked march (variant 8: two time steps per sweep) on an
MI355X: probe series, port samples and DFT planes against the float64
oracle (<= 1e-4 relative L2), against the single-step fused kernel (the
same per-cell arithmetic, <= 1e-6), bit-identical over segmentations and
between graph replay and direct launches, with an odd step count (the
last step on the two-pass kernels), on a grid spanning several tiles, the
mixed case on both lanes and the graded mixed case."""
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


def _outputs(dr, use_graph=True):
    import torch
    dr.reset()
    dr.run(0, dr.n_steps, use_graph=use_graph)
    torch.cuda.synchronize()
    return dict(ts=dr.time_series().cpu().numpy().copy(),
                port=dr.port_samples().cpu().numpy().copy(),
                planes=[a.cpu().numpy().copy() for a in dr.plane_acc],
                fields=[f.cpu().numpy().copy() for f in dr.fields])


def _vs_oracle(out, ref, tol=TOL):
    for p in range(out["ts"].shape[1]):
        e = cases.rel_l2(out["ts"][:, p], ref.time_series[:, p])
        assert e <= tol, ("probe", p, e)
    for q, acc in enumerate(out["planes"]):
        e = cases.rel_l2(acc.view(np.float64), ref.dft_planes[q].view(np.float64))
        assert e <= tol, ("plane", q, e)


def _close(a, b, tol=1e-6):
    assert cases.rel_l2(a["ts"], b["ts"]) <= tol
    for x, y in zip(a["planes"], b["planes"]):
        assert cases.rel_l2(x.view(np.float64), y.view(np.float64)) <= tol
    for x, y in zip(a["fields"], b["fields"]):
        assert cases.rel_l2(x, y) <= tol


def _same(a, b):
    np.testing.assert_array_equal(a["ts"], b["ts"])
    np.testing.assert_array_equal(a["port"], b["port"])
    for x, y in zip(a["planes"], b["planes"]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a["fields"], b["fields"]):
        np.testing.assert_array_equal(x, y)


def test_tb2_tile_case(R):
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    spec = cases.tile_case()
    n = 241
    _say(f"tb2 tile case {n} steps")
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(n, "run")
    g, cfg = o.build_uniform(spec, n)
    assert tuple(g.shape) == tuple(grid.shape) == (57, 47, 92)
    ref = cstep.run_c(cfg, n)
    dr = DeviceRun(prob, n, "cuda")
    at.set_variant(dr, 1, 1)
    fused = _outputs(dr)
    outs = {}
    for seg in (0, 1, 4):
        at.set_variant(dr, 8, seg)
        assert at.get_variant(dr) == (8, seg)
        _say(f"tb2 tile case segments {seg}")
        outs[seg] = _outputs(dr)
        _vs_oracle(outs[seg], ref)
        _close(outs[seg], fused)
    _same(outs[0], outs[1])
    _same(outs[0], outs[4])
    _say("tb2 tile case, direct launches")
    at.set_variant(dr, 8, 0)
    _same(_outputs(dr, use_graph=False), outs[0])
    dr.close()


@pytest.mark.parametrize("lane", ["run", "forward"])
def test_tb2_mixed_case(R, lane):
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    spec = cases.mixed_case(lane)
    n = 301
    _say(f"tb2 mixed case ({lane})")
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(n, lane)
    g, cfg = o.build_uniform(spec, n)
    ref = cstep.run_c(cfg, n)
    dr = DeviceRun(prob, n, "cuda")
    at.set_variant(dr, 8, 0)
    out = _outputs(dr)
    _vs_oracle(out, ref)
    for q in range(len(prob.ports)):
        if np.abs(ref.port_v[q]).max() > 0:
            assert cases.rel_l2(out["port"][:, q, 0], ref.port_v[q]) <= TOL
    at.set_variant(dr, 1, 1)
    _close(out, _outputs(dr))
    dr.close()


def test_tb2_graded_case(R):
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    spec = cases.mixed_graded_case()
    n = 300
    _say("tb2 graded mixed case")
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(n, "run", s11_freqs=np.asarray(spec["s11_freqs"], np.float32))
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, n, want_fields=False)
    dr = DeviceRun(prob, n, "cuda")
    for seg in (0, 3):
        at.set_variant(dr, 8, seg)
        out = _outputs(dr)
        for p in range(out["ts"].shape[1]):
            e = cases.rel_l2(out["ts"][:, p], ref.time_series[:, p])
            print("tb2 graded seg", seg, "probe", p, e)
            assert e <= TOL
        assert cases.rel_l2(out["port"][:, 0, 0], ref.port_v[0]) <= TOL
    dr.close()


def test_tb2_refuses_what_it_does_not_run(R):
    from rfx_amd import _lib
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    _say("tb2 refusals")
    spec = cases.mixed_case("run")
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(20, "run")
    one = DeviceRun(prob, 20, "cuda", two_sets=False)
    with pytest.raises(_lib.RfxHipError):
        at.set_variant(one, 8, 0)
    one.close()
    spec = cases.flux_case()
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(20, "run")
    fl = DeviceRun(prob, 20, "cuda")
    with pytest.raises(_lib.RfxHipError):
        at.set_variant(fl, 8, 0)
    fl.close()
