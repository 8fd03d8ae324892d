"""Batched sweeps: many independent simulations advanced together.

The reference runs a sweep of uneven grids one instance after another
(rfx/batch.py:530-559 ``run_batch``; ``vmap_material_sweep``,
rfx/vmap_sweep.py:871-1052, only batches equal shapes).  Here every
instance keeps its own device state and plan, and ``rfxh_batch_run``
advances all of them with one H launch and one E launch per time step
(include/rfxhip.h), so small instances share the GPU instead of each
paying its own launches.
"""
from __future__ import annotations

import ctypes as C
import itertools
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .api import DFTPlaneProbe, FDTDState, Result, Simulation, _warn_if_nonfinite
from .engine import DeviceRun


class Batch:
    """A fixed set of simulations with their device state, run as one batch."""

    def __init__(self, sims: list[Simulation], n_steps, device=None, sweeps=None):
        if not sims:
            raise ValueError("empty batch")
        steps = [int(n_steps)] * len(sims) if np.isscalar(n_steps) else [int(n) for n in n_steps]
        if len(steps) != len(sims):
            raise ValueError("one step count per simulation")
        self.sims = list(sims)
        self.n_steps = steps
        self.runs = []
        self.grids = []
        self.problems = []
        sweeps = [None] * len(self.sims) if sweeps is None else list(sweeps)
        for sim, n, sw in zip(self.sims, steps, sweeps):
            grid, prob, _ = sim._build_problem(n, "run", material_sweep=sw)
            dev = device if device is not None else sim._device
            self.runs.append(DeviceRun(prob, n, dev, two_sets=False))
            self.grids.append(grid)
            self.problems.append(prob)
        devs = {r.device for r in self.runs}
        if len(devs) != 1:
            raise ValueError("all batch instances must live on one device")
        self.device = devs.pop()
        self.lib = self.runs[0].lib
        plans = (C.c_void_p * len(self.runs))(*[r.plan.value for r in self.runs])
        self.handle = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.rfxh_batch_create(plans, len(self.runs), C.byref(self.handle)))

    def run(self, use_graph: bool = True) -> list[Result]:
        n = len(self.runs)
        step0 = (C.c_int64 * n)(*([0] * n))
        counts = (C.c_int64 * n)(*self.n_steps)
        for r in self.runs:
            r.reset()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _lib.check(self.lib.rfxh_batch_run(self.handle, step0, counts, stream,
                                               1 if use_graph else 0))
        torch.cuda.synchronize(self.device)
        out = []
        for sim, r, g, prob, ns in zip(self.sims, self.runs, self.grids, self.problems,
                                       self.n_steps):
            ts = r.time_series().cpu().numpy().astype(np.float32)
            planes = {e.name: DFTPlaneProbe(a, int(idx), comp, np.asarray(fr), acc.cpu().numpy())
                      for e, (a, idx, comp, fr), acc in zip(sim._dft_planes, prob.planes,
                                                            r.plane_acc)}
            _warn_if_nonfinite(ts, "run_batch")
            out.append(Result(time_series=ts, state=FDTDState(*r.fields, step=ns),
                              dft_planes=planes, grid=g, dt=g.dt,
                              flux_monitors=sim._flux_results(prob, r, prob.flux_areas),
                              ntff_data=sim._ntff_results(prob.ntff_box, r),
                              ntff_box=prob.ntff_box))
        return out

    def close(self):
        if self.handle:
            self.lib.rfxh_batch_destroy(self.handle)
            self.handle = C.c_void_p()
        for r in self.runs:
            r.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_sims(sims: list[Simulation], n_steps, *, use_graph: bool = True,
             device=None) -> list[Result]:
    """Run every simulation for its own step count (an int, or one per
    simulation) as one batch; returns one Result per simulation, in order."""
    b = Batch(sims, n_steps, device)
    try:
        return b.run(use_graph)
    finally:
        b.close()


# ---- the reference's batch entry points ---------------------------------------

class ParameterSweep:
    """A multi-dimensional parameter sweep: the Cartesian product of the value
    lists given as keyword arguments (rfx/batch.py:39-72)."""

    def __init__(self, **params):
        self._keys = list(params.keys())
        self._values = [np.asarray(v).ravel() for v in params.values()]

    @property
    def keys(self) -> list:
        return list(self._keys)

    @property
    def total(self) -> int:
        n = 1
        for v in self._values:
            n *= len(v)
        return n

    def combinations(self):
        """Every combination as a dict of plain Python values, in product order."""
        for combo in itertools.product(*self._values):
            yield dict(zip(self._keys, [c.item() if isinstance(c, np.generic) else c
                                        for c in combo]))


_BATCH_RUN_KWARGS = {"n_steps", "num_periods", "use_graph"}


def run_batch(sim_factory, sweep: ParameterSweep, *, run_kwargs: dict | None = None):
    """Run the simulation ``sim_factory(**params)`` for every combination of
    ``sweep`` with ``sim.run(**run_kwargs)``; returns ``[(params, Result)]``
    in sweep order (rfx/batch.py:530-559).

    The reference runs the combinations one after another.  Here, when
    ``run_kwargs`` asks for nothing beyond ``n_steps`` / ``num_periods`` /
    ``use_graph``, every simulation advances together in one batch (one H
    launch and one E launch per time step covering all instances that still
    have steps left, ``rfxh_batch_run``); other options (``autotune``,
    ``s_param_freqs``, ...) fall back to ``sim.run(**run_kwargs)`` per
    combination."""
    run_kwargs = dict(run_kwargs or {})
    params = list(sweep.combinations())
    sims = [sim_factory(**p) for p in params]
    if not sims:
        return []
    if not set(run_kwargs) <= _BATCH_RUN_KWARGS or not all(isinstance(s, Simulation) for s in sims):
        return [(p, s.run(**run_kwargs)) for p, s in zip(params, sims)]
    n = run_kwargs.get("n_steps")
    if n is None:
        periods = float(run_kwargs.get("num_periods", 20.0))
        steps = [s._build_grid().num_timesteps(periods) for s in sims]
    else:
        steps = [int(n)] * len(sims)
    res = run_sims(sims, steps, use_graph=bool(run_kwargs.get("use_graph", True)))
    return list(zip(params, res))


@dataclass
class VmapSweepResult:
    """Result of a material sweep (rfx/vmap_sweep.py:58-98): ``time_series``
    (n_batch, n_steps, n_probes), the swept name and values, ``dft_planes``
    keyed by plane name, each (n_batch, n_freqs, n1, n2) complex, or None."""
    time_series: np.ndarray
    param_name: str
    param_values: np.ndarray
    final_fields: dict | None = None
    dft_planes: dict | None = None

    def peak_field(self) -> np.ndarray:
        """Peak |probe value| per batch element."""
        return np.max(np.abs(self.time_series), axis=(1, 2))


_SWEEP_FIELDS = ("eps_r", "sigma", "mu_r")


def _parse_param_name(param_name: str):
    """``"eps_r"`` or ``"substrate.eps_r"`` -> (material name or None, field)
    (rfx/vmap_sweep.py:108-125)."""
    if "." in param_name:
        name, fld = param_name.rsplit(".", 1)
    else:
        name, fld = None, param_name
    if fld not in _SWEEP_FIELDS:
        raise ValueError(f"param_name field must be one of {_SWEEP_FIELDS}, got {fld!r} "
                         f"(from {param_name!r})")
    return name, fld


def vmap_material_sweep(sim: Simulation, param_name: str, param_values, *, n_steps=None,
                        num_periods: float = 20.0, return_fields: bool = False
                        ) -> VmapSweepResult:
    """Evaluate ``sim`` for every value of one material parameter
    (rfx/vmap_sweep.py:871-1052).  ``param_name`` is ``"eps_r"``, ``"sigma"``
    or ``"mu_r"`` (every cell whose value differs from vacuum takes the swept
    value, CPML pads included) or ``"<material>.<field>"`` (the cells inside
    that material's shapes take it, then the pads are extended again).

    The reference vmaps its scan over the batched material arrays; here each
    value gets its own material-id table over the same geometry and all of
    them run in one batch (``rfxh_batch_run``: one H and one E launch per
    time step for the whole sweep)."""
    if return_fields:
        raise ValueError("return_fields=True is not implemented on vmap_material_sweep "
                         "(the reference raises too); use sim.run() for a final-field "
                         "snapshot")
    values = np.asarray(param_values, dtype=np.float32).ravel()
    if len(values) == 0:
        raise ValueError("param_values must not be empty")
    name, fld = _parse_param_name(param_name)
    if name is not None and name not in sim._materials:
        raise ValueError(f"unknown material {name!r}")
    if n_steps is None:
        n_steps = sim._build_grid().num_timesteps(num_periods)
    n_steps = int(n_steps)
    b = Batch([sim] * len(values), n_steps,
              sweeps=[(name, fld, float(v)) for v in values])
    try:
        out = b.run()
    finally:
        b.close()
    ts = np.stack([r.time_series for r in out])
    planes = None
    if sim._dft_planes:
        planes = {e.name: np.stack([r.dft_planes[e.name].accumulator for r in out])
                  for e in sim._dft_planes}
    return VmapSweepResult(time_series=ts, param_name=param_name, param_values=values,
                           dft_planes=planes)
