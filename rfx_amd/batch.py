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

import numpy as np
import torch

from . import _lib
from .api import DFTPlaneProbe, FDTDState, Result, Simulation, _warn_if_nonfinite
from .engine import DeviceRun


class Batch:
    """A fixed set of simulations with their device state, run as one batch."""

    def __init__(self, sims: list[Simulation], n_steps, device=None):
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
        for sim, n in zip(self.sims, steps):
            grid, prob, _ = sim._build_problem(n, "run")
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


def run_batch(sims: list[Simulation], n_steps, *, use_graph: bool = True,
              device=None) -> list[Result]:
    """Run every simulation for its own step count (an int, or one per
    simulation); returns one Result per simulation, in order."""
    b = Batch(sims, n_steps, device)
    try:
        return b.run(use_graph)
    finally:
        b.close()
