"""The temporal-blocked march (variant 8: two time steps per sweep) on an
MI355X: probe series, port samples and DFT planes against the float64
oracle (<= 1e-4 relative L2) and against the single-step fused kernel
(the same per-cell arithmetic with its multiply-adds formed in other
places: within 4x the case's own float32 rounding level per output, the
distance of the float32 C oracle from the float64 one), over
segmentations, bit-
identical between graph replay and direct launches, with odd step counts
(the last step on the two-pass kernels), on grids spanning several tiles
(with and without tiles that skip the CPML code, and with PEC walls), the
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


def _f32_levels(cfg, n):
    """Relative L2 distance of the float32 C oracle from the float64 one on
    this case, per output (series, each DFT plane, each final field): the
    rounding level two float32 step kernels that form their multiply-adds
    differently can drift apart by."""
    a = cstep.run_c(cfg, n, double=True)
    b = cstep.run_c(cfg, n, double=False)
    return dict(ts=cases.rel_l2(b.time_series, a.time_series),
                planes=[cases.rel_l2(x.view(np.float64), y.view(np.float64))
                        for x, y in zip(b.dft_planes, a.dft_planes)],
                fields=[cases.rel_l2(b.fields[c], a.fields[c]) for c in o.COMPS])


def _close(a, b, lv):
    """a and b within 4x the float32 rounding level of each output (at
    least 1e-6)."""
    tol = lambda x: max(1e-6, 4.0 * x)   # noqa: E731
    assert cases.rel_l2(a["ts"], b["ts"]) <= tol(lv["ts"])
    for x, y, t in zip(a["planes"], b["planes"], lv["planes"]):
        assert cases.rel_l2(x.view(np.float64), y.view(np.float64)) <= tol(t)
    for c, (x, y, t) in enumerate(zip(a["fields"], b["fields"], lv["fields"])):
        assert cases.rel_l2(x, y) <= tol(t), (o.COMPS[c], cases.rel_l2(x, y), t)


def _same(a, b):
    np.testing.assert_array_equal(a["ts"], b["ts"])
    np.testing.assert_array_equal(a["port"], b["port"])
    for x, y in zip(a["planes"], b["planes"]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a["fields"], b["fields"]):
        np.testing.assert_array_equal(x, y)


def _case_vs_fused(spec, n, segs, shape=None):
    """The sweep at several middle-segment counts vs the oracle (1e-4) and
    vs the fused single step (4x the float32 rounding level of the case,
    at least 1e-6); graph replay vs direct launches bit-identical."""
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(n, "run")
    g, cfg = o.build_uniform(spec, n)
    if shape is not None:
        assert tuple(g.shape) == tuple(grid.shape) == shape
    ref = cstep.run_c(cfg, n)
    lv = _f32_levels(cfg, n)
    dr = DeviceRun(prob, n, "cuda")
    at.set_variant(dr, 1, 1)
    fused = _outputs(dr)
    outs = {}
    for seg in segs:
        at.set_variant(dr, 8, seg)
        assert at.get_variant(dr) == (8, seg)
        _say(f"tb2 {grid.shape} middle segments {seg}")
        outs[seg] = _outputs(dr)
        _vs_oracle(outs[seg], ref)
        _close(outs[seg], fused, lv)
        _close(outs[seg], outs[segs[0]], lv)
    _say(f"tb2 {grid.shape}, direct launches")
    at.set_variant(dr, 8, segs[0])
    _same(_outputs(dr, use_graph=False), outs[segs[0]])
    dr.close()


def test_tb2_tile_case(R):
    _say("tb2 tile case 241 steps")
    _case_vs_fused(cases.tile_case(), 241, (0, 1, 4), (57, 47, 92))


def test_tb2_interior_tiles(R):
    """Tiles that run their middle segments without CPML code beside tiles
    that touch the slabs."""
    _say("tb2 interior-tile case 200 steps")
    _case_vs_fused(cases.tb_interior_case(), 200, (0, 1, 3), (43, 34, 130))


def test_tb2_pec_walls(R):
    """A plan without CPML: every workgroup takes the CPML-free sweep."""
    spec = cases.tb_interior_case()
    spec["boundary"], spec["cpml_layers"] = "pec", 0
    _say("tb2 interior-tile case with PEC walls")
    _case_vs_fused(spec, 160, (0, 2))


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
    _close(out, _outputs(dr), _f32_levels(cfg, n))
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
    # more point operations in one 16 x 64 tile column than the sweep stages
    import rfx_amd as Rx
    sim = Rx.Simulation(10e9, (0.030, 0.010, 0.010), dx=1e-3, boundary="cpml", cpml_layers=2)
    for a in range(1, 30):
        for b in range(1, 10):
            sim.add_source((a * 1e-3, 5e-3, b * 1e-3), "ez", waveform=Rx.GaussianPulse(f0=5e9))
    _, prob, _ = sim._build_problem(20, "run")
    assert len(prob.sources) == 29 * 9
    dense = DeviceRun(prob, 20, "cuda")
    with pytest.raises(_lib.RfxHipError):
        at.set_variant(dense, 8, 0)
    at.set_variant(dense, 4, 0)
    dense.run(0, 20)
    dense.close()
