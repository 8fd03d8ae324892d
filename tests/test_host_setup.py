"""CPU checks of the product's host-side set-up against the oracle's:
material assembly (ids -> eps/sigma/mu), PEC tangential bits, source
tables, probe cells and CPML tables.  No device calls."""
import numpy as np
import pytest

from oracle import fdtd_np as o
from tests import cases


def _setups(spec, n, lane):
    sim = cases.to_sim(spec)
    grid, prob, ports = sim._build_problem(n, lane, spec.get("eps_override"), None, None,
                                           spec.get("s11_freqs"))
    g, cfg = o.build_uniform(spec, n)
    return grid, prob, ports, g, cfg


@pytest.mark.parametrize("name,lane", [("quickstart", "run"), ("mixed", "run"),
                                       ("mixed", "forward"), ("s11", "forward")])
def test_materials_sources_match_oracle(name, lane):
    spec = {"quickstart": cases.quickstart_case, "mixed": lambda: cases.mixed_case(lane),
            "s11": lambda: dict(cases.S11_CASE)}[name]()
    n = 64
    grid, prob, ports, g, cfg = _setups(spec, n, lane)
    assert grid.shape == g.shape and abs(grid.dt - g.dt) == 0.0
    eps, sig, mu = prob.mat.cell_arrays()
    np.testing.assert_array_equal(eps, cfg.eps_r)
    np.testing.assert_array_equal(sig, cfg.sigma)
    np.testing.assert_array_equal(mu, cfg.mu_r)
    # PEC tangential bits
    _, _, _, flg = prob.mat.props()
    bits = flg[prob.mat.ids] & 7
    if cfg.pec_mask is None:
        assert not bits.any()
    else:
        mx, my, mz = o.tangential_edge_masks(cfg.pec_mask)
        np.testing.assert_array_equal(bits & 1, mx.astype(np.uint32))
        np.testing.assert_array_equal((bits >> 1) & 1, my.astype(np.uint32))
        np.testing.assert_array_equal((bits >> 2) & 1, mz.astype(np.uint32))
    assert prob.pec_axes == cfg.pec_axes
    assert len(prob.sources) == len(cfg.sources)
    for (i1, c1, w1), (i2, c2, w2) in zip(prob.sources, cfg.sources):
        assert tuple(i1) == tuple(i2) and c1 == c2
        np.testing.assert_array_equal(w1[:n], w2[:n])
    assert [tuple(p[0]) for p in prob.probes] == [tuple(p[0]) for p in cfg.probes]
    assert [(p[0], p[1]) for p in prob.planes] == [(p[0], p[1]) for p in cfg.dft_planes]
    for f, (n_l, b, c, km1, inv_d) in enumerate(prob.faces):
        ax, side = f // 2, f % 2
        if cfg.cpml_n == 0:
            assert n_l == 0
            continue
        ob, oc, ok = cfg.cpml_profiles[(ax, side)]
        np.testing.assert_array_equal(b, ob)
        np.testing.assert_array_equal(c, oc)
        np.testing.assert_allclose(km1, 1.0 / ok - 1.0, atol=1e-7)


def test_coefficient_table_float32_order():
    """ca, cb, ce, ch follow yee.py:386-393 and cpml.py:538-546/817 in float32."""
    from rfx_amd.materials import MaterialMap
    m = MaterialMap((4, 4, 4))
    m.get_id(2.2, 0.01, 1.0)
    m.get_id(4.4, 0.0, 3.0)
    coef, flg = m.coefficient_table(3.8131e-12)
    eps_r = np.array([1.0, 2.2, 4.4], np.float32)
    sig = np.array([0.0, 0.01, 0.0], np.float32)
    mu = np.array([1.0, 1.0, 3.0], np.float32)
    dt = np.float32(3.8131e-12)
    eps = eps_r * np.float32(o.EPS_0)
    L = sig * dt / (np.float32(2) * eps)
    np.testing.assert_array_equal(coef[:, 0], (1 - L) / (1 + L))
    np.testing.assert_array_equal(coef[:, 1], (dt / eps) / (1 + L))
    np.testing.assert_array_equal(coef[:, 3], dt / (mu * np.float32(o.MU_0)))


def test_eps_derivative_table_matches_finite_difference():
    """d(ca, cb, ce)/d eps_r of the reverse mode vs a float64 central
    difference of the coefficient formulas."""
    from rfx_amd.adjoint import eps_derivative_table
    from rfx_amd.materials import MaterialMap
    dt = 1.9e-12
    m = MaterialMap((2, 2, 2))
    for e, s in ((2.2, 0.01), (4.4, 0.0), (9.0, 0.05), (1.0, 20.0)):
        m.get_id(e, s, 1.0)
    tab = eps_derivative_table(m, dt)
    eps_r, sig, _, _ = m.props()

    def coefs(e, s):
        eps = e * o.EPS_0
        L = s * dt / (2 * eps)
        return np.array([(1 - L) / (1 + L), (dt / eps) / (1 + L), dt / eps])

    for q in range(len(eps_r)):
        e, s = float(eps_r[q]), float(sig[q])
        h = 1e-6 * e
        fd = (coefs(e + h, s) - coefs(e - h, s)) / (2 * h)
        np.testing.assert_allclose(tab[q], fd, rtol=1e-5, atol=1e-30)


@pytest.mark.parametrize("name", ["mixed_graded", "cavity", "mixed_graded_xy", "xy_cavity"])
def test_graded_setup_matches_oracle(name):
    """Graded-z lane: grid metrics, materials on the graded node
    coordinates, current-source tables, port folds / metrics, probes,
    planes and CPML tables equal the oracle's build_graded."""
    spec = {"mixed_graded": cases.mixed_graded_case, "mixed_graded_xy": cases.mixed_graded_xy_case,
            "cavity": lambda: cases.graded_cavity_case()[0],
            "xy_cavity": lambda: cases.graded_xy_cavity_case()[0]}[name]()
    n = 48
    sim = cases.to_sim(spec)
    s11 = None if spec.get("s11_freqs") is None else np.asarray(spec["s11_freqs"], np.float32)
    grid, prob, ports = sim._build_problem(n, "run", s11_freqs=s11)
    g, cfg = o.build_graded(spec, n)
    assert grid.shape == g.shape and grid.dt == g.dt and prob.graded
    inv_e, inv_h = grid.inverse_spacings()
    for a in range(3):
        np.testing.assert_array_equal(inv_e[a], g.inv_e[a])
        np.testing.assert_array_equal(inv_h[a], g.inv_h[a])
        np.testing.assert_array_equal(grid.node_coords()[a], g.node_positions(a))
    eps, sig, mu = prob.mat.cell_arrays()
    np.testing.assert_array_equal(eps, cfg.eps_r)
    np.testing.assert_array_equal(sig, cfg.sigma)
    np.testing.assert_array_equal(mu, cfg.mu_r)
    _, _, _, flg = prob.mat.props()
    bits = flg[prob.mat.ids] & 7
    if cfg.pec_mask is None:
        assert not bits.any()
    else:
        assert cfg.pec_mask.any()
        mx, my, mz = o.tangential_edge_masks(cfg.pec_mask)
        np.testing.assert_array_equal(bits & 1, mx.astype(np.uint32))
        np.testing.assert_array_equal((bits >> 1) & 1, my.astype(np.uint32))
        np.testing.assert_array_equal((bits >> 2) & 1, mz.astype(np.uint32))
    assert prob.pec_axes == "xyz" and prob.record_order == "graded"
    assert len(prob.sources) == len(cfg.sources) and len(prob.sources) > 0
    for (i1, c1, w1), (i2, c2, w2) in zip(prob.sources, cfg.sources):
        assert tuple(i1) == tuple(i2) and c1 == c2
        np.testing.assert_array_equal(w1[:n], w2[:n])
    assert [tuple(p[0]) for p in prob.probes] == [tuple(p[0]) for p in cfg.probes]
    assert [(p[0], p[1]) for p in prob.planes] == [(p[0], p[1]) for p in cfg.dft_planes]
    assert len(ports) == len(cfg.ports_vi)
    for p, q in zip(ports, cfg.ports_vi):
        assert tuple(p[0]) == tuple(q[0]) and p[1] == q[1]
        assert p[2] == q[3] and tuple(p[3]) == tuple(q[4])
    for f, (n_l, b, c, km1, inv_d) in enumerate(prob.faces):
        ax, side = f // 2, f % 2
        if cfg.cpml_n == 0:
            assert n_l == 0
            continue
        ob, oc, ok = cfg.cpml_profiles[(ax, side)]
        np.testing.assert_array_equal(b, ob)
        np.testing.assert_array_equal(c, oc)
        assert inv_d == np.float32(1.0 / cfg.cpml_face_d[ax][side]) or abs(
            inv_d * cfg.cpml_face_d[ax][side] - 1) < 1e-7


@pytest.mark.parametrize("graded", [False, True])
def test_flux_setup_matches_oracle(graded):
    """Flux monitors: plane index, tangential region, cell areas and window
    weights equal the oracle's set-up (uniform: finite boxes and a Tukey
    window; graded: full planes with per-cell widths)."""
    if graded:
        spec = cases.mixed_graded_case()
        spec["flux_monitors"] = [("z", 0.006, None, 3, None, None, "hann", 0.25),
                                 ("x", 0.004, None, 2, None, None, "rect", 0.25)]
        build = o.build_graded
    else:
        spec = cases.flux_case()
        spec["flux_monitors"] = spec["flux_monitors"] + [("y", 0.012, None, 4, None, None,
                                                          "tukey", 0.3)]
        build = o.build_uniform
    n = 64
    sim = cases.to_sim(spec)
    grid, prob, _ = sim._build_problem(n, "run")
    g, cfg = build(spec, n)
    assert len(prob.fluxes) == len(cfg.flux) == len(prob.flux_areas)
    for mine, ref, a1, a2 in zip(prob.fluxes, cfg.flux, prob.flux_areas, cfg.flux_area):
        assert tuple(mine[:6]) == tuple(ref[:6])
        np.testing.assert_array_equal(np.asarray(mine[6], np.float32), ref[6])
        np.testing.assert_array_equal(a1, a2)
        if ref[7] is None:
            assert mine[7] is None
        else:
            np.testing.assert_allclose(mine[7], ref[7], rtol=0, atol=1e-7)
