"""rfx_amd — MI355X (gfx950) backend for rfx's FDTD time-stepping path.

Same user surface as the reference for that path:
``Simulation(...).add_*(...).run(n_steps) -> Result.time_series`` and the
differentiable ``Simulation.forward(eps_override=...)``.  The time loop runs
as hand-written HIP kernels (rfx_amd/csrc/rfxhip.hip) behind the C ABI in
include/rfxhip.h; torch provides device memory, streams and autograd.
"""
from .api import (DFTPlaneProbe, FDTDState, FluxMonitor, ForwardResult, Result, Simulation,
                  extract_lumped_s11, flux_spectrum)
from .boundary import Boundary, BoundarySpec
from .geometry import Box
from .grid import Grid
from .materials import DebyePole, LorentzPole, drude_pole, lorentz_pole
from .sources import GaussianPulse, LumpedPort

__all__ = ["Simulation", "Result", "ForwardResult", "FDTDState", "DFTPlaneProbe", "FluxMonitor",
           "flux_spectrum", "Box",
           "GaussianPulse", "LumpedPort", "Grid", "extract_lumped_s11", "DebyePole",
           "LorentzPole", "lorentz_pole", "drude_pole", "Boundary", "BoundarySpec"]
__version__ = "0.1.0"
