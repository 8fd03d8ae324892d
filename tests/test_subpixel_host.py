"""Subpixel smoothing on the host (no GPU): the Kottke average against its
closed forms at planar interfaces (to the float32 rounding of the node
coordinates, 2e-5 relative: the fill fraction at a node on a face is
1/2, and 0 or 1 half a cell away, only up to that rounding, amplified
by the eps contrast), the oracle's per-component update
reducing to the isotropic one, the product's smoothing equal to the
oracle's, and the per-component material ids."""
import numpy as np
import pytest

from oracle import cstep
from oracle import fdtd_np as o
from rfx_amd import smoothing
from rfx_amd.materials import ANISO_FLAG, MaterialMap
from tests import cases

F32 = np.float32


def _slab_spec(x_lo, x_hi, eps=4.0, dx=1e-3):
    """A dielectric slab spanning y and z (past the domain) between x_lo
    and x_hi, in a 20 x 8 x 8 mm box with 4 CPML layers."""
    return dict(freq_max=20e9, domain=(0.020, 0.008, 0.008), dx=dx, boundary="cpml",
                cpml_layers=4, materials={"slab": (eps, 0.0, 1.0)},
                geometry=[((x_lo, -0.01, -0.01), (x_hi, 0.02, 0.02), "slab")],
                ports=[((0.004, 0.004, 0.004), "ez", 0.0, o.Pulse(10e9, 0.8), True)],
                probes=[((0.015, 0.004, 0.004), "ez")], dft_planes=[], lane="run",
                subpixel=True)


def _grids(spec):
    sim = cases.to_sim(spec)
    grid = sim._build_grid()
    boxes = [(lo, hi, spec["materials"][n][0]) for lo, hi, n in spec["geometry"]]
    g, cfg = o.build_uniform(spec, 4)
    return grid, boxes, cfg


@pytest.mark.parametrize("eps", [4.0, 11.7])
def test_faces_on_nodes_give_the_arithmetic_mean_tangentially(eps):
    """Slab faces on E nodes (x = 8 mm, 12 mm): Ey, Ez sample the face with
    f = 1/2 and the normal along x, so eps_yy = eps_zz = (eps + 1)/2 there
    (tangential: arithmetic mean); Ex samples half a cell off the face are
    bulk (f = 0 or 1)."""
    grid, boxes, cfg = _grids(_slab_spec(0.008, 0.012, eps))
    pad = grid.axis_pads[0]
    ex, ey, ez = smoothing.smoothed_eps(grid, boxes)
    for face in (8, 12):
        i = face + pad
        assert np.allclose(ey[i], (eps + 1.0) / 2.0, rtol=2e-5)
        assert np.allclose(ez[i], (eps + 1.0) / 2.0, rtol=2e-5)
    assert np.allclose(ex[8 + pad], eps, rtol=2e-5) and np.allclose(ex[7 + pad], 1.0, rtol=2e-5)
    assert np.allclose(ex[11 + pad], eps, rtol=2e-5) and np.allclose(ex[12 + pad], 1.0, rtol=2e-5)
    assert np.allclose(ey[10 + pad], eps, rtol=2e-5) and np.allclose(ey[14 + pad], 1.0, rtol=2e-5)
    for a, b in zip((ex, ey, ez), cfg.aniso_eps):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("eps", [4.0, 11.7])
def test_faces_between_nodes_give_the_harmonic_mean_normally(eps):
    """Slab faces half a cell off the nodes: Ex samples sit on the faces
    (f = 1/2, normal along x): eps_xx = 2 eps / (eps + 1) (normal: harmonic
    mean); Ey and Ez samples are bulk."""
    grid, boxes, cfg = _grids(_slab_spec(0.0085, 0.0125, eps))
    pad = grid.axis_pads[0]
    ex, ey, ez = smoothing.smoothed_eps(grid, boxes)
    for i in (8 + pad, 12 + pad):
        assert np.allclose(ex[i], 2.0 * eps / (eps + 1.0), rtol=2e-5)
    assert np.allclose(ey[9 + pad], eps, rtol=2e-5) and np.allclose(ey[8 + pad], 1.0, rtol=2e-5)
    assert np.allclose(ez[12 + pad], eps, rtol=2e-5) and np.allclose(ez[13 + pad], 1.0, rtol=2e-5)
    for a, b in zip((ex, ey, ez), cfg.aniso_eps):
        assert np.array_equal(a, b)


def test_partial_fill_fraction():
    """Face a quarter cell past a node: the Ey samples of that node plane
    are a quarter cell inside (f = 3/4), tangential to the face."""
    eps = 5.0
    grid, boxes, _ = _grids(_slab_spec(0.00775, 0.012, eps))
    pad = grid.axis_pads[0]
    _, ey, _ = smoothing.smoothed_eps(grid, boxes)
    assert np.allclose(ey[8 + pad], 0.75 * eps + 0.25, rtol=2e-5)


def test_equal_components_reduce_to_the_isotropic_update():
    """The oracle with every component's eps equal to eps_r runs the plain
    update bit for bit."""
    spec = cases.mixed_case("run")
    g, cfg = o.build_uniform(spec, 40)
    plain = cstep.run_c(cfg, 40, want_fields=False)
    cfg.aniso_eps = (cfg.eps_r.copy(), cfg.eps_r.copy(), cfg.eps_r.copy())
    same = cstep.run_c(cfg, 40, want_fields=False)
    assert np.array_equal(plain.time_series, same.time_series)
    np_run = o.run_loop(cfg, 40)
    assert np.array_equal(np_run.time_series, plain.time_series)


@pytest.mark.parametrize("which", ["uniform", "graded"])
def test_product_smoothing_equals_the_oracle(which):
    spec = cases.mixed_case("run") if which == "uniform" else cases.mixed_graded_case()
    spec["subpixel"] = True
    sim = cases.to_sim(spec)
    grid = sim._build_grid()
    boxes = [(lo, hi, spec["materials"][n][0]) for lo, hi, n in spec["geometry"]]
    mine = smoothing.smoothed_eps(grid, boxes)
    g, cfg = (o.build_uniform if which == "uniform" else o.build_graded)(spec, 4)
    for a, b in zip(mine, cfg.aniso_eps):
        assert np.array_equal(a, b)
    assert any(np.any(a != cfg.eps_r) for a in mine)


def test_material_ids_carry_the_components():
    rng = np.random.default_rng(1)
    m = MaterialMap((6, 5, 4))
    m.ids[:] = m.get_id(3.0, 0.01, 1.0)
    m.ids[0] = m.get_id(1.0, 0.0, 1.0)
    base = np.array([e[0] for e in m.table], F32)[m.ids]
    comps = [base.copy() for _ in range(3)]
    for c in comps:
        sel = rng.random(c.shape) < 0.2
        c[sel] = rng.choice(np.array([1.5, 2.0, 2.5], F32), sel.sum())
    n = m.apply_aniso(*comps)
    assert n == int(np.sum((comps[0] != base) | (comps[1] != base) | (comps[2] != base)))
    eps_x = np.array([e[0] for e in m.table], F32)
    for mid, v in m.aniso.items():
        eps_x[mid] = v[0]
        assert m.table[mid][3] & ANISO_FLAG
    assert np.array_equal(eps_x[m.ids], comps[0])
    # a later flag edit keeps the components
    cell = tuple(np.argwhere(np.isin(m.ids, list(m.aniso)))[0])
    before = m.aniso[int(m.ids[cell])]
    m.set_cell(cell, add_flags=8)
    assert m.aniso[int(m.ids[cell])] == before
    coef, _ = m.coefficient_table(1e-12)
    rows = m.aniso_table(1e-12)
    assert rows.shape == (len(m.table), 4)
