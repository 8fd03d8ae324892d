"""Reverse mode through flux monitors, NTFF boxes and the far-field transform
on an MI355X.

* rfxh_ntff_far_field_adjoint is the conjugate transpose of the oracle's
  far-field map: Re <G, A x> (A the oracle's float64 transform, x random
  face data) equals Re <A^H G, x> from the adjoint kernel, to 1e-10.
* J = sum |E_theta|^2 + |E_phi|^2 over a few directions of the mixed case
  with an NTFF box around its sources, through the differentiable face
  accumulators and the transform, against float64 oracle central
  differences w.r.t. eps_r.
* J = sum P(f)^2 of a Hann-windowed flux monitor (H averaged over two
  planes) against float64 oracle central differences."""
import sys

import numpy as np
import pytest

from oracle import fdtd_np as o
from tests import cases

pytestmark = pytest.mark.gpu

THETA = np.linspace(0.3, 2.8, 5)
PHI = np.linspace(0.0, 2 * np.pi, 4, endpoint=False)
N_STEPS = 200
REGION = (5, 20, 6, 15, 5, 14)


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


def _spec():
    spec = cases.mixed_case("forward")
    spec["flux_monitors"] = [("x", 0.012, None, 3, (0.008, 0.006), (0.006, 0.005), "hann", 0.25)]
    spec["ntff"] = ((0.002, 0.002, 0.002), (0.014, 0.010, 0.008),
                    np.array([6e9, 9e9], np.float32), 2)
    return spec


def _oracle(spec):
    g, cfg = o.build_uniform(spec, N_STEPS)
    out = o.run_loop(cfg, N_STEPS)
    (lo, hi, fr), = cfg.ntff
    pos, width = o.ntff_axis_tables(g.shape, g.dx, cpml_lo=(g.pad,) * 3)
    et, ep = o.far_field(out.ntff[0], lo, hi, fr, pos, width, THETA, PHI)
    flux = o.flux_spectrum(out.flux[0], cfg.flux_area[0])
    return cfg, float(np.sum(np.abs(et) ** 2 + np.abs(ep) ** 2)), float(np.sum(flux ** 2))


def _fd_check(grad, base, spec, which, J_scale, seed):
    rng = np.random.default_rng(seed)
    i0, i1, j0, j1, k0, k1 = REGION
    for trial in range(3):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            Js.append(_oracle(sp)[which] / J_scale)
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_far_field_adjoint_is_the_transpose_of_the_oracle_map(torch_mod):
    torch = torch_mod
    from rfx_amd import farfield as F
    _say("far-field adjoint vs the oracle transform (dot-product identity)")
    spec = _spec()
    sim = cases.to_sim(spec)
    grid = sim._build_grid()
    box = sim._ntff_setup(grid)
    g, cfg = o.build_uniform(spec, 4)
    (lo, hi, fr), = cfg.ntff
    assert tuple(lo) == box.lo and tuple(hi) == box.hi
    pos, width = o.ntff_axis_tables(g.shape, g.dx, cpml_lo=(g.pad,) * 3)
    rng = np.random.default_rng(21)
    nf = len(fr)
    x = {name: rng.standard_normal((nf,) + shp + (4,)) + 1j * rng.standard_normal((nf,) + shp + (4,))
         for name, shp in zip(o.NTFF_FACES, o.ntff_face_shapes(lo, hi))}
    G = [rng.standard_normal((nf, len(THETA), len(PHI)))
         + 1j * rng.standard_normal((nf, len(THETA), len(PHI))) for _ in range(2)]
    et, ep = o.far_field(x, lo, hi, fr, pos, width, THETA, PHI)
    lhs = float(np.real(np.sum(np.conj(G[0]) * et) + np.sum(np.conj(G[1]) * ep)))
    # the adjoint through torch autograd on differentiable face data
    xt = [torch.tensor(x[name], dtype=torch.complex128, device="cuda", requires_grad=True)
          for name in F.FACES]
    ff = F.compute_far_field(F.NTFFData(*xt), box, grid, THETA, PHI)
    np.testing.assert_allclose(ff.E_theta.detach().cpu().numpy(), et, rtol=1e-10,
                               atol=1e-10 * np.abs(et).max())
    Gt = [torch.tensor(v, dtype=torch.complex128, device="cuda") for v in G]
    J = torch.real(torch.sum(torch.conj(Gt[0]) * ff.E_theta) + torch.sum(torch.conj(Gt[1]) * ff.E_phi))
    J.backward()
    rhs = sum(float(np.real(np.sum(np.conj(t.grad.cpu().numpy()) * x[name])))
              for t, name in zip(xt, F.FACES))
    print("dot product", lhs, rhs, abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # each face separately (a face left out of the objective gets no share)
    for q, name in enumerate(F.FACES):
        part = float(np.real(np.sum(np.conj(xt[q].grad.cpu().numpy()) * x[name])))
        xo = {k: (v if k == name else np.zeros_like(v)) for k, v in x.items()}
        eo, po = o.far_field(xo, lo, hi, fr, pos, width, THETA, PHI)
        ref = float(np.real(np.sum(np.conj(G[0]) * eo) + np.sum(np.conj(G[1]) * po)))
        assert abs(part - ref) <= 1e-10 * abs(lhs), (name, part, ref)


def test_far_field_gradient_vs_oracle_fd(torch_mod):
    torch = torch_mod
    from rfx_amd import farfield as F
    _say("far-field objective directional FD (NTFF box, mixed case)")
    spec = _spec()
    cfg0, J0, _ = _oracle(spec)
    base = cfg0.eps_r.copy()
    sim = cases.to_sim(spec)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=N_STEPS, grad_region=REGION)
    assert isinstance(res.ntff_data.x_lo, torch.Tensor) and res.ntff_data.x_lo.requires_grad
    ff = F.compute_far_field(res.ntff_data, res.ntff_box, res.grid, THETA, PHI)
    J = torch.sum(ff.E_theta.abs() ** 2 + ff.E_phi.abs() ** 2) / J0
    print("J", float(J.detach()), 1.0)
    assert abs(float(J.detach()) - 1.0) <= 1e-3
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0
    _fd_check(grad, base, spec, 1, J0, 31)


def test_flux_gradient_vs_oracle_fd(torch_mod):
    torch = torch_mod
    from rfx_amd.api import flux_spectrum
    _say("flux-monitor objective directional FD (Hann window, mixed case)")
    spec = _spec()
    cfg0, _, P0 = _oracle(spec)
    base = cfg0.eps_r.copy()
    sim = cases.to_sim(spec)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=N_STEPS, grad_region=REGION)
    (mon,) = res.flux_monitors.values()
    P = flux_spectrum(mon)
    assert isinstance(P, torch.Tensor) and P.requires_grad
    J = torch.sum(P ** 2) / P0
    print("J", float(J.detach()), 1.0)
    assert abs(float(J.detach()) - 1.0) <= 1e-3
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0
    _fd_check(grad, base, spec, 2, P0, 41)
