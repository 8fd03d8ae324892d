"""Reverse mode through Debye and Lorentz materials on an MI355X: the
dispersive record (rfxh_run_record_ex) and the ADE's transpose in the
reverse sweep.  On the Debye case (a Debye box overlapped by a plain
dielectric, a lumped port, CPML) and on the same box with a Lorentz pole
and with Debye plus Lorentz poles: directional derivatives of a weighted
sum of squared probe samples w.r.t. eps_r and sigma against float64 oracle
(numpy ADE) central differences, and the 5-segment recompute equal to the
single-segment sweep (the polarisation state is part of each
checkpoint)."""
import sys

import numpy as np
import pytest

from oracle import fdtd_np as o
from tests import cases

pytestmark = pytest.mark.gpu

N_STEPS = 200
REGION = (5, 20, 5, 16, 4, 14)
W0 = 2 * np.pi * 9e9


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


def _spec(kind):
    spec = cases.debye_case(12e-12, lane="forward")
    if kind == "lorentz":
        spec.pop("debye")
        spec["lorentz"] = {"deb": [(W0, 2e9, 1.2 * W0 ** 2)]}
    elif kind == "both":
        spec["lorentz"] = {"deb": [(W0, 2e9, 1.2 * W0 ** 2)]}
    return spec


def _weights(T, n_probe):
    return np.linspace(0.5, 1.5, n_probe)


def _J_oracle(spec, eps, sig):
    sp = dict(spec)
    sp["eps_override"] = eps.astype(np.float32)
    sp["sigma_override"] = sig.astype(np.float32)
    _, cfg = o.build_uniform(sp, N_STEPS)
    ts = o.run_loop(cfg, N_STEPS).time_series
    return float((ts ** 2 * _weights(*ts.shape)).sum())


def _forward_grad(torch, sim, base_e, base_s, segments=None):
    eps = torch.tensor(base_e, requires_grad=True)
    sig = torch.tensor(base_s, requires_grad=True)
    res = sim.forward(eps_override=eps, sigma_override=sig, n_steps=N_STEPS,
                      grad_region=REGION, checkpoint_segments=segments)
    ts = res.time_series.double()
    w = torch.tensor(_weights(*ts.shape), dtype=torch.float64, device=ts.device)
    J = (ts ** 2 * w).sum()
    J.backward()
    return float(J.detach()), eps.grad.numpy().astype(np.float64), sig.grad.numpy().astype(np.float64)


@pytest.mark.parametrize("kind", ["debye", "lorentz", "both"])
def test_dispersive_gradient_directions_vs_oracle_fd(torch_mod, kind):
    torch = torch_mod
    _say(f"dispersive ({kind}) eps / sigma directional FD")
    spec = _spec(kind)
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_uniform(spec, N_STEPS)
    base_e, base_s = cfg0.eps_r.copy(), cfg0.sigma.copy()
    J, g_e, g_s = _forward_grad(torch, sim, base_e, base_s)
    J0 = _J_oracle(spec, base_e, base_s)
    print("J", J, J0)
    assert abs(J - J0) <= 1e-3 * abs(J0)
    assert np.abs(g_e).max() > 0 and np.abs(g_s).max() > 0
    rng = np.random.default_rng(17)
    i0, i1, j0, j1, k0, k1 = REGION
    # sigma steps of 1e-3 S/m: J curves in sigma enough here that 1e-2 leaves
    # a 0.6 % truncation error in the central difference
    for which, grad, h in (("eps", g_e, 1e-3), ("eps", g_e, 1e-3), ("sigma", g_s, 1e-3),
                           ("sigma", g_s, 1e-3)):
        d = np.zeros_like(base_e, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        Js = []
        for sgn in (1, -1):
            e, s = base_e.astype(np.float64), base_s.astype(np.float64)
            if which == "eps":
                e = e + sgn * h * d
            else:
                s = s + sgn * h * d
            Js.append(_J_oracle(spec, e, s))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print(kind, which, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_dispersive_segments_equal_single(torch_mod):
    torch = torch_mod
    _say("dispersive (Debye + Lorentz) gradient: 5 segments vs one")
    spec = _spec("both")
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_uniform(spec, N_STEPS)
    base_e, base_s = cfg0.eps_r.copy(), cfg0.sigma.copy()
    J1, e1, s1 = _forward_grad(torch, sim, base_e, base_s)
    J5, e5, s5 = _forward_grad(torch, sim, base_e, base_s, segments=5)
    assert J5 == J1
    np.testing.assert_allclose(e5, e1, rtol=1e-5, atol=1e-6 * np.abs(e1).max())
    np.testing.assert_allclose(s5, s1, rtol=1e-5, atol=1e-6 * np.abs(s1).max())
