"""Per-axis boundary description accepted by ``Simulation(boundary=...)``
(rfx/boundaries/spec.py:46-200: ``Boundary(lo, hi, lo_thickness,
hi_thickness)`` per axis, ``BoundarySpec(x, y, z)`` taking a token string,
a dict or a Boundary per axis).

The HIP backend runs one boundary kind on all six faces: ``cpml`` (with
one layer count) or ``pec``.  ``uniform()`` reduces a spec to that kind
and layer count and raises NotImplementedError for anything else
(periodic, PMC or UPML faces, mixed kinds, per-face layer counts that
differ).
"""
from __future__ import annotations

from dataclasses import dataclass

TOKENS = ("cpml", "upml", "pec", "pmc", "periodic")


def _token(value) -> str:
    if not isinstance(value, str):
        raise TypeError(f"boundary token must be a string, got {type(value).__name__}")
    t = value.strip().lower()
    if t not in TOKENS:
        raise ValueError(f"unknown boundary token {value!r}; expected one of {TOKENS}")
    return t


@dataclass(frozen=True)
class Boundary:
    """Both faces of one axis (rfx/boundaries/spec.py:46-125)."""
    lo: str
    hi: str
    lo_thickness: int | None = None
    hi_thickness: int | None = None

    def __post_init__(self):
        object.__setattr__(self, "lo", _token(self.lo))
        object.__setattr__(self, "hi", _token(self.hi))
        if (self.lo == "periodic") != (self.hi == "periodic"):
            raise ValueError(f"periodic must be symmetric on an axis; got lo={self.lo!r} "
                             f"hi={self.hi!r}")
        for side, tok, th in (("lo", self.lo, self.lo_thickness), ("hi", self.hi, self.hi_thickness)):
            if th is None:
                continue
            if tok not in ("cpml", "upml"):
                raise ValueError(f"{side}_thickness is only meaningful for absorbing faces, "
                                 f"got {side}={tok!r}")
            if not isinstance(th, int) or th < 0:
                raise ValueError(f"{side}_thickness must be a non-negative int, got {th!r}")

    @classmethod
    def coerce(cls, value) -> "Boundary":
        if isinstance(value, Boundary):
            return value
        if isinstance(value, dict):
            return cls(value["lo"], value["hi"], value.get("lo_thickness"),
                       value.get("hi_thickness"))
        if isinstance(value, str):
            return cls(value, value)
        raise TypeError(f"axis boundary must be a string, Boundary or dict, got "
                        f"{type(value).__name__}")


@dataclass(frozen=True, init=False)
class BoundarySpec:
    """Boundary of each axis (rfx/boundaries/spec.py:156-200)."""
    x: Boundary
    y: Boundary
    z: Boundary

    def __init__(self, x="cpml", y="cpml", z="cpml"):
        object.__setattr__(self, "x", Boundary.coerce(x))
        object.__setattr__(self, "y", Boundary.coerce(y))
        object.__setattr__(self, "z", Boundary.coerce(z))

    def faces(self):
        """(token, thickness or None) of the six faces, x lo first."""
        return [(b.lo, b.lo_thickness) if s == 0 else (b.hi, b.hi_thickness)
                for b in (self.x, self.y, self.z) for s in (0, 1)]

    def uniform(self, cpml_layers: int) -> tuple[str, int]:
        """The one kind on every face and its CPML layer count."""
        faces = self.faces()
        kinds = {t for t, _ in faces}
        if len(kinds) != 1 or next(iter(kinds)) not in ("cpml", "pec"):
            raise NotImplementedError(
                f"the HIP backend runs one boundary kind, 'cpml' or 'pec', on all six faces; "
                f"got {[t for t, _ in faces]}")
        kind = kinds.pop()
        layers = {cpml_layers if th is None else th for _, th in faces}
        if len(layers) != 1:
            raise NotImplementedError(f"per-face CPML layer counts are not supported: {sorted(layers)}")
        return kind, layers.pop()
