"""Batched reverse mode on an MI355X (rfxh_batch_run_record /
rfxh_batch_adjoint): a reduced batched-sweep workload -- four instances of
the sweep's recipe at 8-14 cell interiors with 8-layer CPML and uneven
step counts -- whose d(sum of squared probe samples)/d eps_r equals the
per-instance reverse mode, does not depend on the segment length, and
matches float64 oracle central differences."""
import sys

import numpy as np
import pytest

import bench
from oracle import cstep
from oracle import fdtd_np as o

pytestmark = pytest.mark.gpu

SIZES = [(10, 10, 10), (12, 10, 8), (8, 12, 10), (14, 8, 8)]


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rfx_amd  # noqa: F401
    return torch


def _specs():
    out = []
    for sp, n in bench.c5_specs(4, sizes=SIZES, steps=(120, 221), thick=(2, 5), seed=11):
        # a pulse that reaches the probes within the run
        sp["sources"] = [(pos, comp, 15e9, 0.5) for pos, comp, _, _ in sp["sources"]]
        out.append((sp, n))
    return out


def _problems(specs):
    probs = []
    for sp, n in specs:
        sim = bench.build_sim(sp)
        _, prob, _ = sim._build_problem(n, "forward")
        probs.append(prob)
    return probs


def _batched(torch, specs, segment_steps=None):
    from rfx_amd.adjoint import BatchAdjointRunner
    probs = _problems(specs)
    run = BatchAdjointRunner(probs, [n for _, n in specs], "cuda", segment_steps=segment_steps)
    try:
        ts = run.forward()
        grads = run.backward([2.0 * t for t in ts])
        return ([t.cpu().numpy().copy() for t in ts], [g.cpu().numpy().copy() for g in grads],
                run.S)
    finally:
        run.close()


def test_batched_equals_per_instance(torch_mod):
    torch = torch_mod
    from rfx_amd.adjoint import AdjointRunner
    specs = _specs()
    _say("batched reverse mode vs per-instance, 4 instances")
    ts_b, g_b, S = _batched(torch, specs)
    assert S == max(n for _, n in specs)        # one segment: the forward records
    for q, (prob, (_, n)) in enumerate(zip(_problems(specs), specs)):
        r = AdjointRunner(prob, n, "cuda")
        ts = r.forward()
        g = r.backward(2.0 * ts).cpu().numpy()
        np.testing.assert_allclose(ts_b[q], ts.cpu().numpy(), rtol=1e-6,
                                   atol=1e-6 * np.abs(ts_b[q]).max())
        assert np.abs(g).max() > 0
        np.testing.assert_allclose(g_b[q], g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
        r.dr.close()


def test_batched_segments_equal_single(torch_mod):
    torch = torch_mod
    specs = _specs()
    _say("batched reverse mode: 37-step segments vs one segment")
    ts1, g1, _ = _batched(torch, specs)
    ts2, g2, S = _batched(torch, specs, segment_steps=37)
    assert S == 37
    for q in range(len(specs)):
        np.testing.assert_array_equal(ts2[q], ts1[q])
        np.testing.assert_allclose(g2[q], g1[q], rtol=1e-5, atol=1e-6 * np.abs(g1[q]).max())


def test_batched_gradient_vs_oracle_fd(torch_mod):
    """Directional derivatives of J = sum of squared probe samples for every
    instance vs float64 oracle central differences (eps_r perturbed over
    the whole grid)."""
    torch = torch_mod
    specs = _specs()
    _say("batched reverse mode vs oracle central differences")
    ts_b, g_b, S = _batched(torch, specs, segment_steps=50)
    rng = np.random.default_rng(9)
    for q, (sp, n) in enumerate(specs):
        osp = bench.oracle_spec(sp)
        osp["lane"] = "forward"
        _, cfg0 = o.build_uniform(osp, n)
        base = cfg0.eps_r.copy()
        grad = g_b[q].astype(np.float64)
        d = rng.uniform(-1, 1, base.shape)
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            p2 = dict(osp)
            p2["eps_override"] = (base + sgn * h * d).astype(np.float32)
            _, cfg = o.build_uniform(p2, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(float((out.time_series ** 2).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("instance", q, n, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale
