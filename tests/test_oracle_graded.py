"""Pins of the graded-z oracle lane (oracle/fdtd_np.py build_graded,
GradedGrid, current_source_f32, port metrics) by closed forms and
invariants; each check is chosen so that a plausible mistake fails it, and
a planted copy of that mistake (in the oracle, on the CPU) shows it does."""
import copy

import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from tests import cases
from tests.spectral import matrix_pencil_freqs

F32 = np.float32


def _tm111_error(g, cfg, n, f111):
    out = cstep.run_c(cfg, n, want_fields=False)
    ts = out.time_series[:, 0]
    t_src = 2 * 3.0 / (f111 * 0.8 * np.pi)         # 2 t0 of the pulse: source is off
    n0 = int(np.ceil(t_src / g.dt))
    D = 16
    f, _ = matrix_pencil_freqs(ts[n0::D], g.dt * D)
    fs = f[(f > 0.6 * f111) & (f < 1.5 * f111)]
    by = sorted(fs, key=lambda x: abs(x - f111))
    return abs(by[0] - f111) / f111, abs(by[1] - f111)


def test_graded_cavity_tm111_matches_closed_form():
    """Air PEC cavity, 4:1 graded z: TM111 (z-dependent) within 0.04 % of
    (c/2) sqrt((1/a)^2 + (1/b)^2 + (1/d)^2), the reference's gate
    (tests/test_nonuniform_cavity_accuracy.py:176-200); mode isolated."""
    spec, n, f111, d = cases.graded_cavity_case()
    g, cfg = o.build_graded(spec, n)
    assert max(spec["dz_profile"]) / min(spec["dz_profile"]) > 2.0
    err, second = _tm111_error(g, cfg, n, f111)
    print("TM111 rel err", err)
    assert err < 4e-4
    assert second > 0.5e9


def test_cavity_gate_catches_a_missing_bounding_node():
    """Planted mistake: N cells carried by N nodes (no duplicated bounding
    cell) shortens the realised z extent by one cell; the gate must trip."""
    spec, n, f111, d = cases.graded_cavity_case()
    g, cfg = o.build_graded(spec, n)
    bad = copy.copy(cfg)
    bad.shape = cfg.shape[:2] + (cfg.shape[2] - 1,)
    for k in ("eps_r", "sigma", "mu_r"):
        setattr(bad, k, getattr(cfg, k)[:, :, :-1].copy())
    dz = g.d32[2][:-1]
    inv_h = np.concatenate([(F32(1) / dz)[:-1], np.zeros(1, F32)])
    inv_e = np.concatenate([(F32(1) / dz)[:1], F32(2) / (dz[:-1] + dz[1:])])
    bad.spacing = dict(kind="graded", inv_e=[g.inv_e[0], g.inv_e[1], inv_e],
                       inv_h=[g.inv_h[0], g.inv_h[1], inv_h])
    err, _ = _tm111_error(g, bad, n, f111)
    assert err > 4e-4 * 5


def _stencil_errors(g, inv_e, inv_h):
    """One H update from E linear in node position and one E update from H
    linear in cell-centre position, each along every axis: a consistent
    graded Yee difference is exact for linear fields (first-order
    consistency), so dH = -ch * slope and dE = cb * slope at every cell away
    from the zero-padded ends.  Returns the max relative error."""
    shape = g.shape
    cfg = o.StepConfig(shape=shape, dt=g.dt, spacing=dict(kind="graded", inv_e=inv_e, inv_h=inv_h),
                       eps_r=np.ones(shape, F32), sigma=np.zeros(shape, F32),
                       mu_r=np.ones(shape, F32), pec_axes="")
    worst = 0.0
    # (E component, H component it drives, sign of d/d axis in curl) per axis
    cases_h = {0: ("ez", "hy", +1.0), 1: ("ex", "hz", +1.0), 2: ("ey", "hx", +1.0)}
    cases_e = {0: ("hz", "ey", -1.0), 1: ("hx", "ez", -1.0), 2: ("hy", "ex", -1.0)}
    for ax in range(3):
        d = g.d32[ax].astype(np.float64)
        nodes = np.concatenate([[0.0], np.cumsum(d)])[:-1]
        centres = nodes + d / 2
        shp = [1, 1, 1]
        shp[ax] = -1
        st = o.Stepper(cfg)
        ec, hc, sgn = cases_h[ax]
        st.f[ec] = np.broadcast_to(nodes.reshape(shp), shape).copy()
        st.update_h()
        got = st.f[hc]
        want = -st.ch * sgn * 1.0 * np.ones(shape)
        sl = [slice(1, -1)] * 3
        sl[ax] = slice(0, shape[ax] - 1)
        # curl H_y has -dEz/dx (sign handled by sgn below)
        r = np.abs(np.abs(got[tuple(sl)]) - np.abs(want[tuple(sl)])).max() / np.abs(want).max()
        worst = max(worst, r)
        st = o.Stepper(cfg)
        hcomp, ecomp, sgn = cases_e[ax]
        st.f[hcomp] = np.broadcast_to(centres.reshape(shp), shape).copy()
        st.update_e()
        got = st.f[ecomp]
        want = st.cb * 1.0
        sl = [slice(1, -1)] * 3
        sl[ax] = slice(1, shape[ax])
        r = np.abs(np.abs(got[tuple(sl)]) - np.abs(want[tuple(sl)])).max() / np.abs(want).max()
        worst = max(worst, r)
    return worst


def _graded3():
    prof = cases.smooth_grading([1e-3] * 4 + [0.25e-3] * 4 + [1e-3] * 4, 1.3)
    return o.GradedGrid(prof, 1e-3, None, cpml_layers=2,
                        dx_profile=cases.smooth_grading([1e-3] * 3 + [0.4e-3] * 3 + [1e-3] * 3, 1.3),
                        dy_profile=cases.smooth_grading([1e-3] * 3 + [0.5e-3] * 5, 1.3))


def test_graded_stencil_exact_on_linear_fields():
    g = _graded3()
    assert _stencil_errors(g, g.inv_e, g.inv_h) < 1e-6


def test_stencil_check_catches_local_width_e_metric():
    """Planted mistake: the E update divided by the local cell width instead
    of the mean of the two adjacent widths (the pre-fix stencil)."""
    g = _graded3()
    bad = [F32(1) / d for d in g.d32]
    assert _stencil_errors(g, bad, g.inv_h) > 1e-2


def _radiate(dz, comp, n=900):
    spec = dict(freq_max=20e9, domain=(0.03, 0.03), dx=1e-3, dz_profile=np.asarray(dz),
                boundary="cpml", cpml_layers=8,
                ports=[((0.015, 0.015, 0.015), comp, 0.0, o.Pulse(10e9, 0.8), True)],
                probes=[((0.015, 0.027, 0.015), comp)])
    g, cfg = o.build_graded(spec, n)
    out = cstep.run_c(cfg, n, want_fields=False)
    return g, cfg, out.time_series[:, 0]


@pytest.mark.parametrize("comp", ["ex", "ez"])
def test_current_source_moment_is_mesh_independent(comp):
    """A current source normalised by its node's control volume radiates the
    same field whether its node sits on a uniform mesh or on a 2:1 grading
    step (mixed primal/dual volume, rfx/nonuniform.py:627-679).  The
    broadside co-polarised field agrees to grid dispersion (< 1 % peak);
    the all-primal volume would be 25 % off at the step and fails."""
    uni = [1e-3] * 30
    step = [1e-3] * 10 + [0.5e-3] * 10 + [1e-3] * 15
    g1, c1, t1 = _radiate(uni, comp)
    g2, c2, t2 = _radiate(step, comp)
    k = c2.sources[0][0][2]
    assert g2.d32[2][k - 1] != g2.d32[2][k]          # the node sits on the step
    tt1 = np.arange(len(t1)) * g1.dt
    b = np.interp(tt1, np.arange(len(t2)) * g2.dt, t2)
    ratio = np.abs(b).max() / np.abs(t1).max()
    assert abs(ratio - 1) < 0.01 and np.linalg.norm(b - t1) / np.linalg.norm(t1) < 0.03
    # the same comparison under the all-primal volume
    a = {"ex": 0, "ey": 1, "ez": 2}[comp]
    idx = c2.sources[0][0]
    mixed = np.prod([float(g2.d32[x][idx[x]]) if x == a else g2.dual64(x, idx[x]) for x in range(3)])
    primal = np.prod([float(g2.d32[x][idx[x]]) for x in range(3)])
    if comp != "ez":
        assert abs(ratio * mixed / primal - 1) > 0.2


@pytest.mark.parametrize("comp", ["ex", "ey", "ez"])
@pytest.mark.parametrize("ab", [(1.0, 0.0), (0.0, 1.0)])
def test_port_loop_current_is_exact_for_linear_h(comp, ab):
    """Stokes on a graded dual face: with H_c = a x_c and H_b = b x_b (each
    linear along the other transverse axis, at its own staggered centres),
    the discrete Ampere loop equals (a - b) dual_b dual_c exactly when each
    leg carries the dual spacing of its own axis."""
    g = _graded3()
    a_, b_ = ab
    axis = {"ex": 0, "ey": 1, "ez": 2}[comp]
    ax_b, ax_c = [(1, 2), (2, 0), (0, 1)][axis]       # (a, b, c) cyclic
    names = {0: "hx", 1: "hy", 2: "hz"}
    f = {c: np.zeros(g.shape) for c in ("hx", "hy", "hz")}

    def centres(ax):
        d = g.d32[ax].astype(np.float64)
        return np.concatenate([[0.0], np.cumsum(d)])[:-1] + d / 2

    shp_b = [1, 1, 1]
    shp_b[ax_b] = -1
    shp_c = [1, 1, 1]
    shp_c[ax_c] = -1
    # H_c varies along b, H_b varies along c
    f[names[ax_c]] = np.broadcast_to(a_ * centres(ax_b).reshape(shp_b), g.shape).copy()
    f[names[ax_b]] = np.broadcast_to(b_ * centres(ax_c).reshape(shp_c), g.shape).copy()
    cell = (g.shape[0] // 2, g.shape[1] // 2, g.shape[2] // 2)
    _, _, _, loop = o._graded_port_metrics(g, cell, axis)
    got = o.port_current(f, cell, comp, loop)
    duals = [g.dual64(0, cell[0]), g.dual64(1, cell[1]), g.dual_z32(cell[2])]
    closed = (a_ - b_) * duals[ax_b] * duals[ax_c]
    assert abs(got / closed - 1) < 1e-6


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_port_fold_obeys_ohm_on_the_dual_face(axis):
    """The folded conductivity realises Z0 between the edge's ends: the
    conduction current sigma E through the dual face (dual1 dual2) equals
    V / Z0 with V = E d_par on the primal edge."""
    g = _graded3()
    cell = (g.shape[0] // 2 + 1, g.shape[1] // 2 + 1, g.shape[2] // 2)
    d_par, p1, p2, _ = o._graded_port_metrics(g, cell, axis)
    z0 = 50.0
    sigma = d_par / (z0 * p1 * p2)
    E = 3.0
    duals = [g.dual64(0, cell[0]), g.dual64(1, cell[1]), g.dual_z32(cell[2])]
    area = np.prod([duals[x] for x in range(3) if x != axis])
    prim = [float(g.d32[x][cell[x]]) for x in range(3)][axis]
    assert abs(sigma * E * area - E * prim / z0) <= 1e-6 * E * prim / z0


def test_uniform_profile_reduces_to_uniform_lane():
    """A uniform dz profile on the graded lane reproduces the uniform run()
    lane: same shape, materials, CPML; the current source differs from the
    J source by exactly dV = dx^3 (rfx/nonuniform.py:627-679 vs
    rfx/simulation.py:239-293), and dt by the c used (grid C0 vs
    1/sqrt(eps0 mu0)), which is forced equal here."""
    n = 300
    dx = 1e-3
    spec_u = dict(freq_max=10e9, domain=(0.016, 0.012, 0.010), dx=dx, boundary="cpml",
                  cpml_layers=4, materials={"d": (3.0, 0.02, 1.0), "mag": (1.5, 0.0, 2.0)},
                  geometry=[((0.0045, 0.0035, 0.0025), (0.0095, 0.0085, 0.0065), "d"),
                            ((0.0105, 0.0025, 0.0035), (0.0135, 0.0055, 0.0085), "mag")],
                  ports=[((0.003, 0.006, 0.005), "ez", 0.0, o.Pulse(8e9, 0.8), True)],
                  probes=[((0.013, 0.006, 0.005), "ez"), ((0.010, 0.005, 0.004), "hx")],
                  lane="run")
    gu, cu = o.build_uniform(spec_u, n)
    spec_g = dict(spec_u)
    spec_g["domain"] = spec_u["domain"][:2]
    spec_g["dz_profile"] = np.full(10, dx)
    gg, cg = o.build_graded(spec_g, n)
    assert gg.shape == gu.shape
    np.testing.assert_array_equal(cg.eps_r, cu.eps_r)
    np.testing.assert_array_equal(cg.mu_r, cu.mu_r)
    assert abs(gg.dt / gu.dt - 1) < 1e-8
    cg.dt = cu.dt
    ou = o.run_loop(cu, n)
    og = o.run_loop(cg, n)
    ts_g = og.time_series * dx ** 3
    rel = np.linalg.norm(ts_g - ou.time_series) / np.linalg.norm(ou.time_series)
    assert rel < 1e-5


def _tm110_error(g, cfg, n, f):
    out = cstep.run_c(cfg, n, want_fields=False)
    ts = out.time_series[:, 0]
    n0 = int(np.ceil(2 * 3.0 / (f * 0.8 * np.pi) / g.dt))
    fr, _ = matrix_pencil_freqs(ts[n0::16], g.dt * 16)
    fs = fr[(fr > 0.6 * f) & (fr < 1.5 * f)]
    by = sorted(fs, key=lambda x: abs(x - f))
    return abs(by[0] - f) / f, abs(by[1] - f)


def test_graded_xy_cavity_tm110_matches_closed_form():
    """x- and y-graded air PEC cavity: TM110 within 0.05 % of
    (c/2) sqrt((1/a)^2 + (1/b)^2), the reference's gate
    (tests/test_nonuniform_xy_cavity_accuracy.py:100-125); measured 0.031 %."""
    spec, n, f110 = cases.graded_xy_cavity_case()
    g, cfg = o.build_graded(spec, n)
    err, second = _tm110_error(g, cfg, n, f110)
    print("TM110 rel err", err)
    assert err < 5e-4 and second > 0.3e9


def test_xy_cavity_gate_catches_a_missing_bounding_node_on_x():
    """Planted mistake: the x axis without its duplicated bounding cell."""
    spec, n, f110 = cases.graded_xy_cavity_case()
    g, cfg = o.build_graded(spec, n)
    bad = copy.copy(cfg)
    bad.shape = (cfg.shape[0] - 1,) + cfg.shape[1:]
    for k in ("eps_r", "sigma", "mu_r"):
        setattr(bad, k, getattr(cfg, k)[:-1].copy())
    dxa = g.d32[0][:-1]
    inv_h = np.concatenate([(F32(1) / dxa)[:-1], np.zeros(1, F32)])
    inv_e = np.concatenate([(F32(1) / dxa)[:1], F32(2) / (dxa[:-1] + dxa[1:])])
    bad.spacing = dict(kind="graded", inv_e=[inv_e, g.inv_e[1], g.inv_e[2]],
                       inv_h=[inv_h, g.inv_h[1], g.inv_h[2]])
    bad.sources = [(s[0], s[1], s[2]) for s in cfg.sources]
    err, _ = _tm110_error(g, bad, n, f110)
    assert err > 5e-4 * 5
