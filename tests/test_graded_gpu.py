"""Graded-z lane on an MI355X vs the graded oracle (float64 C restatement):
probe series, DFT planes and the wire-port S matrix to <= 1e-4 relative,
and the TM111 resonance of the graded PEC cavity."""
import sys

import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from tests import cases
from tests.spectral import matrix_pencil_freqs

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


def test_mixed_graded_vs_oracle(R):
    _say("mixed graded case 25x21x27 x 600")
    spec = cases.mixed_graded_case()
    n = 600
    sim = cases.to_sim(spec)
    res = sim.run(n_steps=n, s_param_freqs=spec["s11_freqs"])
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, n)
    for p in range(ref.time_series.shape[1]):
        e = cases.rel_l2(res.time_series[:, p], ref.time_series[:, p])
        print("probe", p, e)
        assert e <= TOL
    for q, name in enumerate(res.dft_planes):
        acc = res.dft_planes[name].accumulator
        e = cases.rel_l2(acc.view(np.float64), ref.dft_planes[q].view(np.float64))
        print("plane", name, e)
        assert e <= TOL
    s_ref = o.extract_lumped_s11(ref.port_v_dft[0], ref.port_i_dft[0], 50.0)
    s = res.s_params[0, 0]
    print("S11", s, s_ref)
    np.testing.assert_allclose(s, s_ref, rtol=TOL)


def test_graded_cavity_vs_oracle_and_tm111(R):
    _say("graded PEC cavity 41x36x53 x 8000")
    spec, n, f111, d = cases.graded_cavity_case()
    sim = cases.to_sim(spec)
    res = sim.run(n_steps=n)
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, 2000, want_fields=False)
    e = cases.rel_l2(res.time_series[:2000], ref.time_series)
    print("cavity first 2000 steps rel-L2", e)
    assert e <= TOL
    ts = res.time_series[:, 0].astype(np.float64)
    n0 = int(np.ceil(2 * 3.0 / (f111 * 0.8 * np.pi) / g.dt))
    f, _ = matrix_pencil_freqs(ts[n0::16], g.dt * 16, rel_tol=1e-6)
    fs = f[(f > 0.6 * f111) & (f < 1.5 * f111)]
    err = np.min(np.abs(fs - f111)) / f111
    print("TM111 rel err (float32 run)", err)
    assert err < 4e-4


def test_mixed_graded_xy_vs_oracle(R):
    _say("mixed case graded in x, y and z")
    spec = cases.mixed_graded_xy_case()
    n = 500
    sim = cases.to_sim(spec)
    res = sim.run(n_steps=n, s_param_freqs=spec["s11_freqs"])
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, n)
    for p in range(ref.time_series.shape[1]):
        e = cases.rel_l2(res.time_series[:, p], ref.time_series[:, p])
        print("probe", p, e)
        assert e <= TOL
    for q, name in enumerate(res.dft_planes):
        e = cases.rel_l2(res.dft_planes[name].accumulator.view(np.float64),
                         ref.dft_planes[q].view(np.float64))
        assert e <= TOL
    s_ref = o.extract_lumped_s11(ref.port_v_dft[0], ref.port_i_dft[0], 50.0)
    np.testing.assert_allclose(res.s_params[0, 0], s_ref, rtol=TOL)


def test_graded_xy_cavity_tm110(R):
    _say("x/y-graded PEC cavity x 8000 (TM110)")
    spec, n, f110 = cases.graded_xy_cavity_case()
    sim = cases.to_sim(spec)
    res = sim.run(n_steps=n)
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, 2000, want_fields=False)
    assert cases.rel_l2(res.time_series[:2000], ref.time_series) <= TOL
    ts = res.time_series[:, 0].astype(np.float64)
    n0 = int(np.ceil(2 * 3.0 / (f110 * 0.8 * np.pi) / g.dt))
    f, _ = matrix_pencil_freqs(ts[n0::16], g.dt * 16, rel_tol=1e-6)
    fs = f[(f > 0.6 * f110) & (f < 1.5 * f110)]
    err = np.min(np.abs(fs - f110)) / f110
    print("TM110 rel err (float32 run)", err)
    assert err < 5e-4


def test_patch_workload_vs_oracle(R):
    """The graded patch workload (PEC ground and patch sheets, FR4, wire
    port through the substrate) at 60 x 60 interior cells."""
    _say("graded patch 85x85x145 x 600")
    spec = cases.patch_case(60)
    n = 600
    sim = cases.to_sim(spec)
    res = sim.run(n_steps=n, s_param_freqs=spec["s11_freqs"])
    g, cfg = o.build_graded(spec, n)
    ref = cstep.run_c(cfg, n, want_fields=False)
    e = cases.rel_l2(res.time_series, ref.time_series)
    print("patch probe rel-L2", e)
    assert e <= TOL
    s_ref = o.extract_lumped_s11(ref.port_v_dft[0], ref.port_i_dft[0], 50.0)
    err = np.abs(res.s_params[0, 0] - s_ref) / np.abs(s_ref)
    print("patch S11 max rel err", err.max())
    assert err.max() <= 1e-3
