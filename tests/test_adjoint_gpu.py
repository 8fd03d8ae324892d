"""Reverse mode on an MI355X: transpose identity, finite differences, the
gradient-ladder fixture, segmented recomputation."""
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
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rfx_amd  # noqa: F401
    return torch


@pytest.mark.parametrize("lane", ["forward", "run"])
def test_adjoint_is_transpose_of_step(torch_mod, lane):
    """<F x0, g> == <x0, F^T g> for random initial fields/CPML memory and
    random probe weights, no sources (the step is then linear)."""
    torch = torch_mod
    _say(f"adjoint dot-product test ({lane} lane)")
    import ctypes as C
    from rfx_amd import _lib
    from rfx_amd.adjoint import AdjointRunner
    spec = cases.mixed_case(lane)
    spec["ports"] = []
    spec["s11_freqs"] = None
    spec["probes"] = spec["probes"] + [((0.012, 0.003, 0.007), "ey"), ((0.002, 0.010, 0.002), "hy")]
    sim = cases.to_sim(spec)
    n = 60
    _, prob, _ = sim._build_problem(n, lane)
    run = AdjointRunner(prob, n, "cuda")
    dr = run.dr
    g = torch.Generator(device="cpu").manual_seed(7)
    x0 = [torch.randn(prob.shape, generator=g).cuda() for _ in range(6)]
    psi0 = torch.randn(dr.psis[0].numel(), generator=g).cuda()
    dr.load_state(x0, psi0)
    ts = run.forward()
    w = torch.randn(ts.shape, generator=g).cuda()
    lhs = float((ts.double() * w.double()).sum())
    # reverse sweep with the probe weights as dJ/dy and no gradient output
    T = n
    probe_adj = w.float().contiguous()
    npsi = psi0.numel()
    adj = [torch.zeros(prob.shape, device="cuda") for _ in range(6)]
    adj_psi = torch.zeros(npsi, device="cuda")
    gbar = torch.zeros(npsi, device="cuda")
    d = _lib.AdjointDesc()
    for c in range(6):
        d.d_adj[c] = adj[c].data_ptr()
    d.d_adj_psi = adj_psi.data_ptr()
    d.d_gbar = gbar.data_ptr()
    d.d_probe_adj = probe_adj.data_ptr()
    _lib.check(dr.lib.rfxh_adjoint(dr.plan, C.byref(d), 0, T, dr.stream_ptr()))
    torch.cuda.synchronize()
    rhs = sum(float((a.double() * x.double()).sum()) for a, x in zip(adj, x0))
    rhs += float((adj_psi.double() * psi0.double()).sum())
    # float32 rounding scales with |F x0| |g|, not with the inner product
    scale = float(ts.double().norm() * w.double().norm())
    print("dot test", lhs, rhs, abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-5 * scale


def _ladder_J(torch, sim, eps, n, dt, region=None, segments=None):
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region,
                      checkpoint_segments=segments)
    ts = res.time_series[:, 0].double()
    t = torch.arange(n, dtype=torch.float64, device=ts.device) * dt
    ph = torch.exp(-1j * 2 * np.pi * 8e9 * t)
    acc = (ts.to(torch.complex128) * ph).sum()
    return (acc.real ** 2 + acc.imag ** 2), res


@pytest.mark.parametrize("rung", [0, 1, 2, 3])
def test_ladder_gradient_vs_fixture_and_fd(torch_mod, rung):
    """Slab-eps gradient on every rung of the reference's gradient ladder
    (lambda/20 .. lambda/60) vs its jax.grad value and the float64 oracle's
    central difference (<= 1e-3)."""
    torch = torch_mod
    lad = cases.ladder()["rungs"][rung]
    N = lad["lambda_over_dx"]
    _say(f"ladder rung lambda/{N} gradient")
    spec, n, m1d = cases.ladder_case(N, 4.0)
    assert n == lad["n_steps"]
    sim = cases.to_sim(spec)
    eps = torch.tensor(spec["eps_override"], requires_grad=True)
    J, res = _ladder_J(torch, sim, eps, n, sim._build_grid().dt)
    J.backward()
    g = eps.grad.numpy()
    d_slab = float(g[m1d].sum())
    # float64 oracle central difference, step 1e-3
    h = 1e-3
    vals = []
    for e in (4.0 + h, 4.0 - h):
        sp, _, _ = cases.ladder_case(N, e)
        gg, cfg = o.build_uniform(sp, n)
        out = cstep.run_c(cfg, n, want_fields=False)
        vals.append(cases.ladder_objective(out.time_series[:, 0], gg.dt))
    fd = (vals[0] - vals[1]) / (2 * h)
    print("grad slab", d_slab, "oracle FD", fd, "jax.grad", lad["grad_ad"])
    assert abs(d_slab - fd) / abs(fd) < 1e-3
    assert abs(d_slab - lad["grad_ad"]) / abs(lad["grad_ad"]) < 1e-3


def test_gradient_directions_vs_oracle_fd(torch_mod):
    """Mixed lossy/PEC/CPML case, gradient region inside and across the CPML:
    directional derivatives vs float64 oracle central differences."""
    torch = torch_mod
    _say("mixed-case directional FD")
    spec = cases.mixed_case("forward")
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 160
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_uniform(spec, n)
    base = cfg0.eps_r.copy()
    region = (2, 14, 3, 17, 2, 10)   # crosses the 4-layer CPML on x_lo, y_lo, z_lo
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region)
    ts = res.time_series.double()
    wts = torch.linspace(0.5, 1.5, ts.shape[1], dtype=torch.float64, device=ts.device)
    J = (ts ** 2 * wts).sum()
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    rng = np.random.default_rng(3)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(3):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            gg, cfg = o.build_uniform(sp, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(float((out.time_series ** 2 * np.linspace(0.5, 1.5, ts.shape[1])).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        # a random direction can nearly cancel; float32 rounding of the
        # gradient scales with sum |g_i d_i|, not with the projection
        scale = float(np.abs(grad * d).sum())
        print("direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_sigma_gradient_directions_vs_oracle_fd(torch_mod):
    """dJ/d sigma_override on the mixed lossy/PEC/CPML case, region across
    the CPML: directional derivatives vs float64 oracle central
    differences; the eps gradient of the same sweep equals the eps-only
    sweep's."""
    torch = torch_mod
    _say("mixed-case sigma directional FD")
    spec = cases.mixed_case("forward")
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 160
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_uniform(spec, n)
    base_e, base_s = cfg0.eps_r.copy(), cfg0.sigma.copy()
    region = (2, 14, 3, 17, 2, 10)

    def J_of(ts):
        ts = ts.double()
        wts = torch.linspace(0.5, 1.5, ts.shape[1], dtype=torch.float64, device=ts.device)
        return (ts ** 2 * wts).sum()

    eps = torch.tensor(base_e, requires_grad=True)
    sig = torch.tensor(base_s, requires_grad=True)
    res = sim.forward(eps_override=eps, sigma_override=sig, n_steps=n, grad_region=region)
    J_of(res.time_series).backward()
    grad_s = sig.grad.numpy().astype(np.float64)
    eps2 = torch.tensor(base_e, requires_grad=True)
    res2 = sim.forward(eps_override=eps2, sigma_override=torch.tensor(base_s), n_steps=n,
                       grad_region=region)
    J_of(res2.time_series).backward()
    np.testing.assert_allclose(eps.grad.numpy(), eps2.grad.numpy(), rtol=1e-6,
                               atol=1e-7 * float(np.abs(eps2.grad.numpy()).max()))
    assert np.abs(grad_s).max() > 0
    rng = np.random.default_rng(5)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(3):
        d = np.zeros_like(base_s, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-2
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = base_e
            sp["sigma_override"] = (base_s + sgn * h * d).astype(np.float32)
            gg, cfg = o.build_uniform(sp, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(float((out.time_series ** 2 * np.linspace(0.5, 1.5, out.time_series.shape[1])).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad_s * d).sum())
        scale = float(np.abs(grad_s * d).sum())
        print("sigma direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_mu_gradient_directions_vs_oracle_fd(torch_mod):
    """dJ/d mu_r_override on the mixed case (a magnetic box, CPML), region
    across the CPML, from a sweep that also takes dJ/d eps_r and dJ/d sigma:
    directional derivatives vs float64 oracle central differences; the eps
    gradient equals the eps-only sweep's."""
    torch = torch_mod
    _say("mixed-case mu_r directional FD")
    spec = cases.mixed_case("forward")
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 160
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_uniform(spec, n)
    base_e, base_s, base_m = cfg0.eps_r.copy(), cfg0.sigma.copy(), cfg0.mu_r.copy()
    region = (2, 14, 3, 17, 2, 10)

    def J_of(ts):
        ts = ts.double()
        wts = torch.linspace(0.5, 1.5, ts.shape[1], dtype=torch.float64, device=ts.device)
        return (ts ** 2 * wts).sum()

    eps = torch.tensor(base_e, requires_grad=True)
    sig = torch.tensor(base_s, requires_grad=True)
    mu = torch.tensor(base_m, requires_grad=True)
    res = sim.forward(eps_override=eps, sigma_override=sig, mu_r_override=mu, n_steps=n,
                      grad_region=region)
    J_of(res.time_series).backward()
    grad_m = mu.grad.numpy().astype(np.float64)
    eps2 = torch.tensor(base_e, requires_grad=True)
    res2 = sim.forward(eps_override=eps2, sigma_override=torch.tensor(base_s),
                       mu_r_override=torch.tensor(base_m), n_steps=n, grad_region=region)
    J_of(res2.time_series).backward()
    np.testing.assert_allclose(eps.grad.numpy(), eps2.grad.numpy(), rtol=1e-6,
                               atol=1e-7 * float(np.abs(eps2.grad.numpy()).max()))
    assert np.abs(grad_m).max() > 0
    rng = np.random.default_rng(6)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(3):
        d = np.zeros_like(base_m, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = base_e
            sp["sigma_override"] = base_s
            sp["mu_r_override"] = (base_m + sgn * h * d).astype(np.float32)
            gg, cfg = o.build_uniform(sp, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(float((out.time_series ** 2 * np.linspace(0.5, 1.5, out.time_series.shape[1])).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad_m * d).sum())
        scale = float(np.abs(grad_m * d).sum())
        print("mu direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_s11_gradient_directions_vs_oracle_fd(torch_mod):
    """J = sum |S11(f)|^2 of the mixed case's lumped ports (forward lane,
    port V/I samples through the adjoint): directional derivatives w.r.t.
    eps_r vs float64 oracle central differences."""
    torch = torch_mod
    _say("mixed-case S11 objective directional FD")
    spec = cases.mixed_case("forward")
    assert any(p[2] != 0.0 for p in spec["ports"]) and spec.get("s11_freqs") is not None
    n = 200
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_uniform(spec, n)
    base = cfg0.eps_r.copy()
    region = (2, 14, 3, 17, 2, 10)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region,
                      port_s11_freqs=np.asarray(spec["s11_freqs"], np.float32))
    s = res.s_params
    assert isinstance(s, torch.Tensor) and s.requires_grad
    J = (s.abs() ** 2).sum()
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0

    def J_oracle(sp):
        gg, cfg = o.build_uniform(sp, n)
        out = cstep.run_c(cfg, n, want_fields=False)
        tot = 0.0
        for q in range(len(out.port_v_dft)):     # the case's one port, Z0 = 50
            tot += float(np.sum(np.abs(o.extract_lumped_s11(out.port_v_dft[q],
                                                            out.port_i_dft[q])) ** 2))
        return tot

    sp0 = dict(spec)
    sp0["eps_override"] = base
    J0 = J_oracle(sp0)
    print("J", float(J), J0)
    assert abs(float(J) - J0) <= 1e-3 * abs(J0)
    rng = np.random.default_rng(8)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(3):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            Js.append(J_oracle(sp))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("s11 direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_dft_plane_gradient_directions_vs_oracle_fd(torch_mod):
    """J = sum_q sum |acc_q|^2 / S_q over the mixed case's two DFT planes (an
    E_z plane and an H_y plane; S_q the oracle's value at the base point),
    through the differentiable plane accumulators: directional derivatives
    w.r.t. eps_r vs float64 oracle central differences."""
    torch = torch_mod
    _say("mixed-case DFT-plane objective directional FD")
    spec = cases.mixed_case("forward")
    n = 200
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_uniform(spec, n)
    base = cfg0.eps_r.copy()
    S = [float(np.sum(np.abs(a) ** 2)) for a in cstep.run_c(cfg0, n, want_fields=False).dft_planes]
    assert len(S) == 2 and min(S) > 0
    region = (2, 14, 3, 17, 2, 10)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region)
    accs = [res.dft_planes[k].accumulator for k in res.dft_planes]
    assert all(isinstance(a, torch.Tensor) and a.requires_grad for a in accs)
    J = sum((a.abs() ** 2).sum() / s for a, s in zip(accs, S))
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0

    def J_oracle(sp):
        _, cfg = o.build_uniform(sp, n)
        out = cstep.run_c(cfg, n, want_fields=False)
        return sum(float(np.sum(np.abs(a) ** 2)) / s for a, s in zip(out.dft_planes, S))

    print("J", float(J), 2.0)
    assert abs(float(J) - 2.0) <= 1e-3 * 2.0
    rng = np.random.default_rng(12)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(3):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            Js.append(J_oracle(sp))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("plane direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_dft_plane_gradient_single_plane_and_segments(torch_mod):
    """A gradient through one plane only (the other's gradient None) and
    through 7 checkpoint segments equals the one-segment gradient; the H_y
    plane's part and the E_z plane's part add up to the gradient of their
    sum."""
    torch = torch_mod
    _say("DFT-plane gradient: single planes, segments")
    spec = cases.mixed_case("forward")
    n = 160
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_uniform(spec, n)
    base = cfg0.eps_r.copy()

    def grad_of(which, segs=None):
        eps = torch.tensor(base, requires_grad=True)
        res = sim.forward(eps_override=eps, n_steps=n, checkpoint_segments=segs)
        accs = [res.dft_planes[k].accumulator for k in res.dft_planes]
        J = sum((accs[q].abs() ** 2).sum() / (accs[q].detach().abs() ** 2).sum() for q in which)
        J.backward()
        return eps.grad.numpy().copy()

    g_e, g_h, g_both = grad_of([0]), grad_of([1]), grad_of([0, 1])
    assert np.abs(g_e).max() > 0 and np.abs(g_h).max() > 0
    np.testing.assert_allclose(g_e + g_h, g_both, rtol=1e-5, atol=1e-5 * np.abs(g_both).max())
    g_seg = grad_of([0, 1], segs=7)
    np.testing.assert_allclose(g_seg, g_both, rtol=1e-5, atol=1e-5 * np.abs(g_both).max())


def test_graded_dft_plane_gradient_vs_oracle_fd(torch_mod):
    """Graded-z lane: J = sum_q sum |acc_q|^2 / S_q over the case's E_z and
    H_x DFT planes (sampled at t = n dt on this lane), directional
    derivatives w.r.t. eps_r vs float64 oracle central differences."""
    torch = torch_mod
    _say("graded DFT-plane objective directional FD")
    spec = cases.mixed_graded_case()
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 200
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_graded(spec, n)
    base = cfg0.eps_r.copy()
    S = [float(np.sum(np.abs(a) ** 2)) for a in cstep.run_c(cfg0, n, want_fields=False).dft_planes]
    assert len(S) == 2 and min(S) > 0
    region = (3, 15, 4, 14, 2, 12)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region)
    accs = [res.dft_planes[k].accumulator for k in res.dft_planes]
    J = sum((a.abs() ** 2).sum() / s for a, s in zip(accs, S))
    print("J", float(J.detach()), 2.0)
    assert abs(float(J.detach()) - 2.0) <= 2e-3
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0
    rng = np.random.default_rng(13)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(2):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            _, cfg = o.build_graded(sp, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(sum(float(np.sum(np.abs(a) ** 2)) / s for a, s in zip(out.dft_planes, S)))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("graded plane direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_pec_mask_override_forward_and_gradient(torch_mod):
    """forward(pec_mask_override=...) merges hard PEC cells with the
    geometry's (a one-cell sheet and a small block here): the time series
    equals the float64 oracle's with the same mask to 1e-4, and the eps_r
    gradient matches oracle central differences."""
    torch = torch_mod
    _say("pec_mask_override: forward parity and directional FD")
    spec = cases.mixed_case("forward")
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 160
    g0, cfg0 = o.build_uniform(spec, n)
    mask = np.zeros(cfg0.eps_r.shape, bool)
    mask[10, 4:16, 3:14] = True                  # a sheet normal to x
    mask[15:18, 11:13, 8:10] = True              # a small block
    spec["pec_mask_override"] = mask
    _, cfg = o.build_uniform(spec, n)
    assert cfg.pec_mask is not None and cfg.pec_mask[10, 5, 5]
    ref = cstep.run_c(cfg, n, want_fields=False).time_series
    base = cfg.eps_r.copy()
    sim = cases.to_sim(spec)
    eps = torch.tensor(base, requires_grad=True)
    region = (2, 14, 3, 17, 2, 10)
    res = sim.forward(eps_override=eps, pec_mask_override=mask, n_steps=n, grad_region=region)
    ts = res.time_series
    err = cases.rel_l2(ts.detach().cpu().numpy(), ref)
    print("series vs oracle", err)
    assert err <= 1e-4
    no_mask = sim.forward(eps_override=torch.tensor(base), n_steps=n).time_series
    assert cases.rel_l2(no_mask.cpu().numpy(), ref) > 1e-2     # the mask matters here
    J = (ts.double() ** 2).sum()
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    rng = np.random.default_rng(19)
    i0, i1, j0, j1, k0, k1 = region
    d = np.zeros_like(base, dtype=np.float64)
    d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
    h = 1e-3
    Js = []
    for sgn in (1, -1):
        sp = dict(spec)
        sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
        _, c2 = o.build_uniform(sp, n)
        Js.append(float((cstep.run_c(c2, n, want_fields=False).time_series ** 2).sum()))
    fd = (Js[0] - Js[1]) / (2 * h)
    ad = float((grad * d).sum())
    scale = float(np.abs(grad * d).sum())
    print("pec-mask direction", ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
    assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_graded_forward_options(torch_mod):
    """The graded lane's forward options (rfx/nonuniform.py:1589-1804):
    n_warmup leaves the outputs unchanged and, with the gradient region
    outside the fields' reach during the warm-up, the gradient too;
    checkpoint_every = 37 equals one segment; emit_time_series=False gives
    no probe series; the uniform lane refuses all three."""
    torch = torch_mod
    _say("graded forward options: n_warmup, checkpoint_every, emit_time_series")
    spec = cases.mixed_graded_case()
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 160
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_graded(spec, n)
    base = cfg0.eps_r.copy()
    region = (17, 21, 4, 14, 2, 12)    # >= 6 cells from the source at x index 10

    def run(**kw):
        eps = torch.tensor(base, requires_grad=True)
        res = sim.forward(eps_override=eps, n_steps=n, grad_region=region, **kw)
        accs = [res.dft_planes[k].accumulator for k in res.dft_planes]
        J = (res.time_series.double() ** 2).sum() if res.time_series.shape[1] else 0.0
        J = J + sum((a.abs() ** 2).sum() / (a.detach().abs() ** 2).sum() for a in accs)
        J.backward()
        return res, eps.grad.numpy().copy()

    r0, g0 = run()
    r1, g1 = run(n_warmup=2)        # the fields reach at most 4 cells in 2 steps
    np.testing.assert_array_equal(r1.time_series.detach().cpu().numpy(),
                                  r0.time_series.detach().cpu().numpy())
    assert np.abs(g0).max() > 0
    np.testing.assert_allclose(g1, g0, rtol=1e-6, atol=1e-7 * np.abs(g0).max())
    r2, g2 = run(checkpoint_every=37)
    np.testing.assert_allclose(g2, g0, rtol=1e-5, atol=1e-6 * np.abs(g0).max())
    r3, g3 = run(emit_time_series=False)
    assert tuple(r3.time_series.shape) == (n, 0)
    usim = cases.to_sim(cases.mixed_case("forward"))
    for kw in ({"n_warmup": 3}, {"checkpoint_every": 10}, {"emit_time_series": False}):
        with pytest.raises(NotImplementedError):
            usim.forward(n_steps=20, **kw)


def test_graded_s11_gradient_vs_oracle_fd(torch_mod):
    """Graded-z lane with its single-cell and wire ports (V/I sampled after
    the injection on this lane): J = sum |S|^2 over the ports and
    frequencies, directional derivatives w.r.t. eps_r vs float64 oracle
    central differences."""
    torch = torch_mod
    _say("graded S-parameter objective directional FD")
    spec = cases.mixed_graded_case()
    n = 200
    sim = cases.to_sim(spec)
    _, cfg0 = o.build_graded(spec, n)
    base = cfg0.eps_r.copy()
    z0 = [p[2] for p in spec["ports"] if p[2] != 0.0]
    region = (3, 15, 4, 14, 2, 12)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region,
                      port_s11_freqs=np.asarray(spec["s11_freqs"], np.float32))
    s = res.s_params
    assert isinstance(s, torch.Tensor) and s.requires_grad
    J = (s.abs() ** 2).sum()

    def J_oracle(sp):
        _, cfg = o.build_graded(sp, n)
        out = cstep.run_c(cfg, n, want_fields=False)
        return sum(float(np.sum(np.abs(o.extract_lumped_s11(out.port_v_dft[q], out.port_i_dft[q],
                                                             z0[q])) ** 2))
                   for q in range(len(out.port_v_dft)))

    sp0 = dict(spec)
    sp0["eps_override"] = base
    J0 = J_oracle(sp0)
    print("J", float(J.detach()), J0)
    assert abs(float(J.detach()) - J0) <= 1e-3 * abs(J0)
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    assert np.abs(grad).max() > 0
    rng = np.random.default_rng(23)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(2):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            Js.append(J_oracle(sp))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("graded s11 direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale


def test_segmented_gradient_equals_single(torch_mod):
    torch = torch_mod
    _say("checkpoint segments == single segment")
    spec, n, m1d = cases.ladder_case(20, 4.0)
    sim = cases.to_sim(spec)
    grads = []
    for segs in (None, 7):
        eps = torch.tensor(spec["eps_override"], requires_grad=True)
        J, _ = _ladder_J(torch, sim, eps, n, sim._build_grid().dt, segments=segs)
        J.backward()
        grads.append(eps.grad.numpy().copy())
    np.testing.assert_allclose(grads[1], grads[0], rtol=1e-5,
                               atol=1e-6 * np.abs(grads[0]).max())


def test_graded_gradient_directions_vs_oracle_fd(torch_mod):
    """Graded-z lane (CPML, lossy substrate, PEC sheet): directional
    derivatives of a probe-energy objective vs float64 oracle central
    differences."""
    torch = torch_mod
    _say("graded directional FD")
    spec = cases.mixed_graded_case()
    spec["ports"] = [p for p in spec["ports"] if p[2] == 0.0]
    spec["s11_freqs"] = None
    n = 200
    sim = cases.to_sim(spec)
    g0, cfg0 = o.build_graded(spec, n)
    base = cfg0.eps_r.copy()
    region = (3, 15, 4, 14, 2, 12)
    eps = torch.tensor(base, requires_grad=True)
    res = sim.forward(eps_override=eps, n_steps=n, grad_region=region)
    ts = res.time_series.double()
    J = (ts ** 2).sum()
    J.backward()
    grad = eps.grad.numpy().astype(np.float64)
    rng = np.random.default_rng(11)
    i0, i1, j0, j1, k0, k1 = region
    for trial in range(2):
        d = np.zeros_like(base, dtype=np.float64)
        d[i0:i1, j0:j1, k0:k1] = rng.uniform(-1, 1, (i1 - i0, j1 - j0, k1 - k0))
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            sp = dict(spec)
            sp["eps_override"] = (base + sgn * h * d).astype(np.float32)
            gg, cfg = o.build_graded(sp, n)
            out = cstep.run_c(cfg, n, want_fields=False)
            Js.append(float((out.time_series ** 2).sum()))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float((grad * d).sum())
        scale = float(np.abs(grad * d).sum())
        print("graded direction", trial, ad, fd, abs(ad - fd) / abs(fd), abs(ad - fd) / scale)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-4 * scale
