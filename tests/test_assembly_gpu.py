"""Material ids assembled on an MI355X (box fill, CPML pads, port cells,
tangential PEC bits) equal the host assembly's bit for bit -- the id array
and the material table -- on the 533^3 throughput workload, the graded
patch workload with its PEC sheets and wire port, and the small mixed
cases (PEC box, lumped ports, a second mu_r)."""
import sys

import numpy as np
import pytest

import bench
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


def _both(sim, n_steps, lane="run", **kw):
    out = {}
    for how in ("host", "device"):
        sim.material_assembly = how
        grid, prob, _ = sim._build_problem(n_steps, lane, **kw)
        out[how] = prob
    sim.material_assembly = "auto"
    return out["host"], out["device"]


def _same(h, d):
    assert h.mat.ids.shape == d.mat.ids.shape
    assert h.mat.table == d.mat.table
    assert np.array_equal(h.mat.ids, d.mat.ids)
    assert [s[0] for s in h.sources] == [s[0] for s in d.sources]
    for a, b in zip(h.sources, d.sources):
        assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("n_int", [96, 512])
def test_throughput_workload(R, n_int):
    spec = bench.c2_spec(n_int)
    sim = bench.build_sim(spec)
    _say(f"device assembly, 64-box workload at {n_int}^3 interior")
    h, d = _both(sim, 8)
    _same(h, d)
    assert len(h.mat.table) == 65


def test_graded_patch_workload(R):
    spec = bench.c3_spec()
    sim = bench.build_c3_sim(spec)
    _say("device assembly, graded patch workload (PEC sheets, wire port)")
    h, d = _both(sim, 8, s11_freqs=np.asarray(spec["s_freqs"], np.float32))
    _same(h, d)
    # the PEC sheets carry tangential-zero bits
    flags = np.array([t[3] for t in h.mat.table])
    assert np.any(flags[h.mat.ids] & 7)


@pytest.mark.parametrize("lane", ["run", "forward"])
def test_mixed_cases(R, lane):
    _say(f"device assembly, mixed cases ({lane})")
    for spec in (cases.mixed_case(lane), cases.mixed_graded_case()):
        sim = cases.to_sim(spec)
        kw = {}
        if spec.get("s11_freqs") is not None:
            kw["s11_freqs"] = np.asarray(spec["s11_freqs"], np.float32)
        h, d = _both(sim, 8, lane=lane, **kw)
        _same(h, d)


def test_device_assembled_run_equals_host(R):
    """A run from device-assembled ids is the run from host ids."""
    spec = cases.mixed_case("run")
    sim = cases.to_sim(spec)
    sim.material_assembly = "host"
    a = sim.run(n_steps=120).time_series
    sim.material_assembly = "device"
    b = sim.run(n_steps=120).time_series
    assert np.array_equal(a, b)
