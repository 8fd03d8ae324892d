"""Per-axis boundary description accepted by ``Simulation(boundary=...)``.

Same interface as the reference's rfx/boundaries/spec.py:46-300:
``Boundary(lo, hi, lo_thickness, hi_thickness, conformal)`` per axis,
``BoundarySpec(x, y, z)`` taking a token string, a dict or a Boundary per
axis, the ``BoundarySpec.uniform(token)`` constructor, ``faces()`` as
(axis, side, token) triples, ``to_dict`` / ``from_dict`` and
``normalize_boundary``.

The HIP backend runs one boundary kind on all six faces: ``cpml`` (with
one layer count) or ``pec``.  ``run_kind(spec, cpml_layers)`` reduces a
spec to that kind and layer count and raises NotImplementedError for
anything else (periodic, PMC or UPML faces, mixed kinds, per-face layer
counts that differ, conformal PEC axes).
"""
from __future__ import annotations

from dataclasses import dataclass

TOKENS = ("cpml", "upml", "pec", "pmc", "periodic")
ABSORBING = ("cpml", "upml")


def _token(value) -> str:
    if isinstance(value, Boundary):
        raise TypeError("expected a token string, got a Boundary; pass the Boundary to the "
                        "axis argument instead")
    if not isinstance(value, str):
        raise TypeError(f"boundary token must be a string, got {type(value).__name__}")
    t = value.strip().lower()
    if t not in TOKENS:
        raise ValueError(f"unknown boundary token {value!r}; expected one of {TOKENS}")
    return t


@dataclass(frozen=True)
class Boundary:
    """Both faces of one axis (rfx/boundaries/spec.py:46-144)."""
    lo: str
    hi: str
    lo_thickness: int | None = None
    hi_thickness: int | None = None
    conformal: bool = False

    def __post_init__(self):
        object.__setattr__(self, "lo", _token(self.lo))
        object.__setattr__(self, "hi", _token(self.hi))
        if (self.lo == "periodic") != (self.hi == "periodic"):
            raise ValueError(f"periodic must be symmetric on an axis; got lo={self.lo!r} "
                             f"hi={self.hi!r}")
        for side, tok, th in (("lo", self.lo, self.lo_thickness), ("hi", self.hi, self.hi_thickness)):
            if th is None:
                continue
            if tok not in ABSORBING:
                raise ValueError(f"{side}_thickness is only meaningful for absorbing faces, "
                                 f"got {side}={tok!r}")
            if not isinstance(th, int) or th < 0:
                raise ValueError(f"{side}_thickness must be a non-negative int, got {th!r}")
        if self.conformal and (self.lo != "pec" or self.hi != "pec"):
            raise ValueError(f"conformal=True needs 'pec' on both faces; got lo={self.lo!r} "
                             f"hi={self.hi!r}")

    @classmethod
    def from_string(cls, token: str) -> "Boundary":
        t = _token(token)
        return cls(t, t)

    @classmethod
    def from_dict(cls, d: dict) -> "Boundary":
        return cls(d["lo"], d["hi"], d.get("lo_thickness"), d.get("hi_thickness"),
                   bool(d.get("conformal", False)))

    def to_dict(self) -> dict:
        out = {"lo": self.lo, "hi": self.hi}
        if self.lo_thickness is not None:
            out["lo_thickness"] = self.lo_thickness
        if self.hi_thickness is not None:
            out["hi_thickness"] = self.hi_thickness
        if self.conformal:
            out["conformal"] = True
        return out

    def resolved_lo_thickness(self, default: int) -> int:
        return default if self.lo_thickness is None else self.lo_thickness

    def resolved_hi_thickness(self, default: int) -> int:
        return default if self.hi_thickness is None else self.hi_thickness

    @classmethod
    def coerce(cls, value) -> "Boundary":
        if isinstance(value, Boundary):
            return value
        if isinstance(value, dict):
            return cls.from_dict(value)
        if isinstance(value, str):
            return cls.from_string(value)
        raise TypeError(f"axis boundary must be a string, Boundary or dict, got "
                        f"{type(value).__name__}")


@dataclass(frozen=True, init=False)
class BoundarySpec:
    """Boundary of each axis (rfx/boundaries/spec.py:156-272)."""
    x: Boundary
    y: Boundary
    z: Boundary

    def __init__(self, x, y, z):
        object.__setattr__(self, "x", Boundary.coerce(x))
        object.__setattr__(self, "y", Boundary.coerce(y))
        object.__setattr__(self, "z", Boundary.coerce(z))
        absorbing = {t for _, _, t in self.faces()} & set(ABSORBING)
        if len(absorbing) > 1:
            raise ValueError(f"CPML and UPML cannot be mixed on one simulation; got "
                             f"{sorted(absorbing)}")

    @classmethod
    def uniform(cls, token: str) -> "BoundarySpec":
        """The same token on all six faces."""
        b = Boundary.from_string(token)
        return cls(b, b, b)

    @classmethod
    def from_dict(cls, d: dict) -> "BoundarySpec":
        return cls(d["x"], d["y"], d["z"])

    def to_dict(self) -> dict:
        return {"x": self.x.to_dict(), "y": self.y.to_dict(), "z": self.z.to_dict()}

    def faces(self):
        """(axis name, 'lo' | 'hi', token) of the six faces, x lo first."""
        for name, b in (("x", self.x), ("y", self.y), ("z", self.z)):
            yield name, "lo", b.lo
            yield name, "hi", b.hi

    @property
    def absorber_type(self) -> str | None:
        for _, _, t in self.faces():
            if t in ABSORBING:
                return t
        return None

    def periodic_axes(self) -> str:
        return "".join(n for n, b in (("x", self.x), ("y", self.y), ("z", self.z))
                       if b.lo == "periodic")

    def _faces_with(self, token: str) -> set:
        return {f"{a}_{s}" for a, s, t in self.faces() if t == token}

    def pec_faces(self) -> set:
        return self._faces_with("pec")

    def pmc_faces(self) -> set:
        return self._faces_with("pmc")

    def conformal_faces(self) -> set:
        return {f"{n}_{s}" for n, b in (("x", self.x), ("y", self.y), ("z", self.z))
                if b.conformal for s in ("lo", "hi")}


def normalize_boundary(obj, default: str = "cpml") -> BoundarySpec:
    """None / token / dict / BoundarySpec -> BoundarySpec (rfx/boundaries/spec.py:276-300)."""
    if obj is None:
        return BoundarySpec.uniform(default)
    if isinstance(obj, BoundarySpec):
        return obj
    if isinstance(obj, str):
        return BoundarySpec.uniform(obj)
    if isinstance(obj, dict):
        return BoundarySpec.from_dict(obj)
    raise TypeError(f"cannot normalise a boundary of type {type(obj).__name__}")


def run_kind(spec: BoundarySpec, cpml_layers: int) -> tuple[str, int]:
    """The one boundary kind on every face ('cpml' or 'pec') and the CPML
    layer count the backend runs; NotImplementedError for anything else."""
    axes = (spec.x, spec.y, spec.z)
    if any(b.conformal for b in axes):
        raise NotImplementedError("conformal PEC faces (Dey-Mittra face shift) are not supported "
                                  "by the HIP backend")
    tokens = [t for _, _, t in spec.faces()]
    kinds = set(tokens)
    if len(kinds) != 1 or tokens[0] not in ("cpml", "pec"):
        raise NotImplementedError(
            f"the HIP backend runs one boundary kind, 'cpml' or 'pec', on all six faces; "
            f"got {tokens}")
    kind = tokens[0]
    layers = {th for b in axes for th in (b.resolved_lo_thickness(cpml_layers),
                                          b.resolved_hi_thickness(cpml_layers))}
    if len(layers) != 1:
        raise NotImplementedError(f"per-face CPML layer counts are not supported: {sorted(layers)}")
    return kind, layers.pop()
