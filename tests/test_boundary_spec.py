"""Simulation(boundary=BoundarySpec(...)) (CPU): a spec with one kind on all
six faces builds the same problem as the string form, per-face layer
counts that agree set the CPML depth, and every spec the HIP backend does
not run (periodic, PMC, UPML, mixed kinds, unequal layer counts) raises
NotImplementedError; the per-axis validation follows
rfx/boundaries/spec.py:46-125 (symmetric periodic, thickness only on
absorbing faces)."""
import numpy as np
import pytest

from rfx_amd import Boundary, BoundarySpec, Box, Simulation


def _problem(boundary, **kw):
    sim = Simulation(10e9, (0.012, 0.010, 0.008), dx=0.001, boundary=boundary, **kw)
    sim.add_material("d", eps_r=3.0)
    sim.add(Box((0.002, 0.002, 0.002), (0.006, 0.005, 0.004)), material="d")
    sim.add_probe((0.006, 0.005, 0.004), "ez")
    return sim._build_problem(8, "run")


@pytest.mark.parametrize("kind", ["cpml", "pec"])
def test_uniform_spec_equals_string(kind):
    g0, p0, _ = _problem(kind, cpml_layers=6)
    g1, p1, _ = _problem(BoundarySpec(kind, kind, kind), cpml_layers=6)
    assert p0.shape == p1.shape and p0.dt == p1.dt
    np.testing.assert_array_equal(p0.mat.ids, p1.mat.ids)


def test_face_thickness_sets_the_depth():
    th = {"lo": "cpml", "hi": "cpml", "lo_thickness": 5, "hi_thickness": 5}
    g1, p1, _ = _problem(BoundarySpec(th, th, Boundary("cpml", "cpml", 5, 5)), cpml_layers=12)
    g0, p0, _ = _problem("cpml", cpml_layers=5)
    assert p0.shape == p1.shape


@pytest.mark.parametrize("spec", [
    BoundarySpec("periodic", "cpml", "cpml"),
    BoundarySpec("cpml", "cpml", "pmc"),
    BoundarySpec("upml", "upml", "upml"),
    BoundarySpec("cpml", "pec", "cpml"),
    BoundarySpec(Boundary("cpml", "cpml", 4, 6), "cpml", "cpml"),
])
def test_unsupported_specs_raise(spec):
    with pytest.raises(NotImplementedError):
        Simulation(10e9, (0.012, 0.010, 0.008), dx=0.001, boundary=spec)


def test_axis_validation():
    with pytest.raises(ValueError):
        Boundary("periodic", "cpml")
    with pytest.raises(ValueError):
        Boundary("pec", "pec", lo_thickness=4)
    with pytest.raises(ValueError):
        BoundarySpec("absorbing", "cpml", "cpml")
    with pytest.raises(TypeError):
        BoundarySpec(3, "cpml", "cpml")
    assert BoundarySpec(" CPML ", "cpml", "cpml").x == Boundary("cpml", "cpml")


def test_reference_constructors_and_views():
    # the reference's classmethod constructor and (axis, side, token) faces
    spec = BoundarySpec.uniform("pec")
    assert list(spec.faces()) == [(a, s, "pec") for a in "xyz" for s in ("lo", "hi")]
    assert spec.pec_faces() == {f"{a}_{s}" for a in "xyz" for s in ("lo", "hi")}
    assert spec.absorber_type is None
    assert BoundarySpec.uniform("cpml").absorber_type == "cpml"
    d = BoundarySpec("cpml", Boundary("cpml", "cpml", 5, 5), "cpml").to_dict()
    assert BoundarySpec.from_dict(d) == BoundarySpec("cpml", Boundary("cpml", "cpml", 5, 5), "cpml")
    from rfx_amd.boundary import normalize_boundary
    assert normalize_boundary(None) == BoundarySpec.uniform("cpml")
    assert normalize_boundary("pec") == spec
    # BoundarySpec.uniform('pec') runs the PEC lane like the string form
    g0, p0, _ = _problem("pec")
    g1, p1, _ = _problem(BoundarySpec.uniform("pec"))
    assert p0.shape == p1.shape and p0.faces == p1.faces


def test_conformal_axis_is_validated_and_refused():
    with pytest.raises(ValueError):
        Boundary("cpml", "cpml", conformal=True)
    spec = BoundarySpec(Boundary("pec", "pec", conformal=True), "pec", "pec")
    assert spec.conformal_faces() == {"x_lo", "x_hi"}
    with pytest.raises(NotImplementedError):
        Simulation(10e9, (0.012, 0.010, 0.008), dx=0.001, boundary=spec)
    with pytest.raises(ValueError):
        BoundarySpec("cpml", "upml", "cpml")
