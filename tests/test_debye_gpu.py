"""Debye dispersive materials on an MI355X vs the float64 oracle: probe
series and fields to 1e-4 relative L2 (uniform and graded lanes), the
two-pass variants and the batched step bit-identical, and the
fast-relaxation limit equal to a plain dielectric of eps_inf + d_eps."""
import sys

import numpy as np
import pytest

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


def test_debye_uniform_vs_oracle(R):
    _say("Debye case 25x21x19 x 400 (uniform)")
    spec = cases.debye_case(20e-12)
    n = 400
    res = cases.to_sim(spec).run(n_steps=n)
    g, cfg = o.build_uniform(spec, n)
    ref = o.run_loop(cfg, n)
    for p in range(res.time_series.shape[1]):
        e = cases.rel_l2(res.time_series[:, p], ref.time_series[:, p])
        print("probe", p, e)
        assert e <= TOL
    n_f = 100
    res_f = cases.to_sim(spec).run(n_steps=n_f)
    g_f, cfg_f = o.build_uniform(spec, n_f)
    ref_f = o.run_loop(cfg_f, n_f)
    for c in o.COMPS:
        e = cases.rel_l2(getattr(res_f.state, c).cpu().numpy(), ref_f.fields[c])
        print("field", c, e)
        assert e <= TOL


def test_debye_graded_vs_oracle(R):
    _say("Debye substrate on the graded-z mixed case x 300")
    spec = cases.mixed_graded_case()
    spec["debye"] = {"sub": [(2.0, 15e-12), (0.5, 60e-12)]}
    n = 300
    res = cases.to_sim(spec).run(n_steps=n)
    g, cfg = o.build_graded(spec, n)
    assert len(cfg.debye_poles) == 2
    ref = o.run_loop(cfg, n)
    for p in range(res.time_series.shape[1]):
        e = cases.rel_l2(res.time_series[:, p], ref.time_series[:, p])
        print("probe", p, e)
        assert e <= TOL


def test_debye_variants_and_batch_bit_identical(R):
    _say("Debye: column, one-cell and batched steps")
    import torch
    from rfx_amd import autotune as at
    from rfx_amd.batch import run_sims
    from rfx_amd.engine import DeviceRun
    spec = cases.debye_case(20e-12)
    n = 120
    sim = cases.to_sim(spec)
    _, prob, _ = sim._build_problem(n, "run")
    outs = []
    for v in ((4, 1), (4, 0), (0, 1), (6, 1)):
        dr = DeviceRun(prob, n)
        at.set_variant(dr, *v)
        dr.run(0, n)
        torch.cuda.synchronize()
        outs.append([f.cpu().numpy().copy() for f in dr.fields])
        with pytest.raises(Exception):
            at.set_variant(dr, 1, 1)      # the ping-pong steps have no Debye ADE
        dr.close()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            np.testing.assert_array_equal(a, b)
    res = run_sims([cases.to_sim(spec), cases.to_sim(cases.mixed_case("run"))], [n, 80])
    single = cases.to_sim(spec).run(n_steps=n, autotune=False)
    np.testing.assert_array_equal(res[0].time_series, single.time_series)


def test_debye_fast_relaxation_equals_static_dielectric(R):
    _say("Debye tau -> 0 vs plain eps_inf + d_eps")
    dt = o.UniformGrid(10e9, (0.016, 0.012, 0.010), 1e-3, 4).dt
    deb = cases.debye_case(1e-6 * dt, overlap=False)
    plain = cases.debye_case(1e-6 * dt, overlap=False)
    plain.pop("debye")
    plain["materials"]["deb"] = (5.0, 0.01, 1.0)
    n = 200
    a = cases.to_sim(deb).run(n_steps=n).time_series
    b = cases.to_sim(plain).run(n_steps=n).time_series
    e = cases.rel_l2(a, b)
    print("fast limit", e)
    assert e <= TOL


def test_lorentz_cavity_polaritons_gpu(R):
    _say("Lorentz-filled PEC cavity 21x17x4 x 6000")
    from tests.spectral import matrix_pencil_freqs
    spec, n, pred = cases.lorentz_cavity_case()
    res = cases.to_sim(spec).run(n_steps=n)
    g, cfg = o.build_uniform(spec, 1000)
    ref = o.run_loop(cfg, 1000)
    e = cases.rel_l2(res.time_series[:1000, 0], ref.time_series[:, 0])
    print("series (1000 steps)", e)
    assert e <= TOL
    f, _ = matrix_pencil_freqs(res.time_series[:, 0].astype(np.float64), res.dt, max_order=60)
    err = [np.min(np.abs(f - p)) / p for p in pred]
    print("polaritons", pred, err)
    assert max(err) < 1e-5


def test_debye_and_lorentz_vs_oracle(R):
    _say("Debye + Lorentz + Drude case x 400")
    spec = cases.debye_case(20e-12)
    w0 = 2 * np.pi * 9e9
    spec["lorentz"] = {"deb": [(w0, 1e9, 1.2 * w0 ** 2)], "d": [(0.0, 5e9, 4e20)]}
    n = 400
    res = cases.to_sim(spec).run(n_steps=n)
    g, cfg = o.build_uniform(spec, n)
    ref = o.run_loop(cfg, n)
    for p in range(res.time_series.shape[1]):
        e = cases.rel_l2(res.time_series[:, p], ref.time_series[:, p])
        print("probe", p, e)
        assert e <= TOL
