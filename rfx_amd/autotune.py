"""Kernel-variant autotuner.

The plan's step can run as the two-pass kernels (H pass, E pass, in place;
one cell per thread or columns of 2, 3 or 4 cells along x) or, with a
second field set, as one fused kernel (flat or XCD-band
workgroup order), as the barrier-free wave-march kernel (each wavefront
walks a pair of rows along x) or as the marching kernel (6- or 14-row tiles
of the y-z plane walking along x), the last two in a chosen number of x
segments, or as the temporal-blocked march, which advances two time steps
per sweep (include/rfxhip.h ``rfxh_set_variant``).  All
variants compute the same step.  Which is fastest depends on the grid (CPML
share, plane size, cache fit), so the tuner runs a few real steps of the
actual problem under each variant, timed with events on the stream the run
uses, keeps the fastest and resets the state.  Results are cached per
(shape, graded, CPML depths, device) for the process and, optionally, in a
JSON file.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import torch

from . import _lib

_CACHE: dict = {}


def _key(dr) -> str:
    """Everything that decides which variants a plan accepts and how fast
    they run: shape, lane, CPML depths, field sets, dispersion poles, whether
    the material table fits the LDS-staged variants, and the device."""
    p = dr.prob
    dev = torch.cuda.get_device_name(dr.device)
    faces = tuple(int(f[0]) for f in p.faces)
    n_pole = len(p.poles) + len(p.lorentz)
    big = int(len(dr.flags) > 1024)
    extras = f"{len(p.planes)}|{len(p.fluxes)}|{len(p.ntff)}"
    return f"{tuple(p.shape)}|{int(p.graded)}|{faces}|{len(dr.sets)}|{n_pole}|{big}|{extras}|{dev}"


def variants(dr):
    """(kernel, parameter) candidates: two-pass kernels 0 (one cell per
    thread) and 4 / 6 / 5 (columns of 2 / 3 / 4 cells) with workgroup order
    1 (flat) or 0 (XCD bands); with two field sets also the fused step 1
    (same orders) and the marching steps 2 / 3 (6- / 14-row tiles) and the
    wave-march 7 with their number of x segments (0 = automatic), and the
    temporal-blocked march 8 (two time steps per sweep: temporal-block depth
    2 against depth 1 for every other variant) with its x segments."""
    out = [(0, 1), (0, 0)]
    for kern in (4, 6, 5):                  # columns of 2, 3, 4 cells
        out += [(kern, order) for order in (1, 0)]
    nx = dr.prob.shape[0]
    if len(dr.sets) == 2:
        out += [(1, 1), (1, 0)]
        # the wave-march and temporal-blocked steps stage the material table
        # in LDS (at most 1024 ids)
        small = len(dr.flags) <= 1024
        for kern in (2, 3) + ((7,) if small else ()):
            out += [(kern, 0)] + [(kern, s) for s in (4, 8, 16, 32, 64) if s <= max(1, nx // 4)]
        if small:
            out += [(8, 0)] + [(8, s) for s in (4, 8, 10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 24, 32)
                               if s <= max(1, nx // 16)]
    return out


MARCH_ROWS = {2: 6, 3: 14}
ORDERED = {0: "two-pass", 1: "fused", 4: "col2", 5: "col4", 6: "col3"}


def variant_name(v) -> str:
    """two-pass-flat, col2-bands, col3-flat, fused-flat, march14-16, wmarch-8, ..."""
    k, prm = v
    if k == 8:
        return "tb2" if prm == 0 else f"tb2-{prm}"
    if k == 7:
        return "wmarch" if prm == 0 else f"wmarch-{prm}"
    if k in MARCH_ROWS:
        base = f"march{MARCH_ROWS[k]}"
        return base if prm == 0 else f"{base}-{prm}"
    return f"{ORDERED[k]}-{'flat' if prm else 'bands'}"


def parse_variant(name: str):
    if name.startswith("tb2"):
        _, _, seg = name.partition("-")
        return (8, int(seg) if seg else 0)
    if name.startswith("wmarch"):
        _, _, seg = name.partition("-")
        return (7, int(seg) if seg else 0)
    if name.startswith("march"):
        head, _, seg = name.partition("-")
        rows = int(head[5:])
        kern = {r: k for k, r in MARCH_ROWS.items()}[rows]
        return (kern, int(seg) if seg else 0)
    head, _, order = name.rpartition("-")
    kern = {v: k for k, v in ORDERED.items()}[head]
    return (kern, 1 if order == "flat" else 0)


def get_variant(dr):
    f, o = C.c_int(), C.c_int()
    _lib.check(dr.lib.rfxh_get_variant(dr.plan, C.byref(f), C.byref(o)))
    return f.value, o.value


def set_variant(dr, fused: int, order: int):
    _lib.check(dr.lib.rfxh_set_variant(dr.plan, int(fused), int(order)))


def autotune(dr, trial_steps: int = 32, repeats: int = 2, cache_file: str | None = None):
    """Pick the fastest variant for ``dr`` (an engine.DeviceRun); leaves the
    run reset (zero state) with that variant selected.  Returns
    {"variant": (fused, order), "ms_per_step": {variant: ms}}."""
    key = _key(dr)
    if cache_file and not _CACHE and os.path.exists(cache_file):
        with open(cache_file) as fh:
            _CACHE.update({k: tuple(v) for k, v in json.load(fh).items()})
    if key in _CACHE:
        try:
            set_variant(dr, *_CACHE[key])
            dr.reset()
            return {"variant": _CACHE[key], "ms_per_step": None, "cached": True}
        except _lib.RfxHipError:
            del _CACHE[key]               # a stale entry this plan refuses: tune again
    steps = max(1, min(int(trial_steps), dr.n_steps))
    stream = torch.cuda.current_stream(dr.device)
    timings = {}
    for v in variants(dr):
        try:
            set_variant(dr, *v)
        except _lib.RfxHipError:
            continue                      # not available for this plan (e.g. Debye poles)
        dr.reset()
        dr.run(0, steps)                      # warm-up, builds the graph
        best = None
        for _ in range(repeats):
            dr.reset()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            dr.run(0, steps)
            b.record(stream)
            b.synchronize()
            ms = a.elapsed_time(b) / steps
            best = ms if best is None else min(best, ms)
        timings[v] = best
    choice = min(timings, key=timings.get)
    set_variant(dr, *choice)
    dr.reset()
    _CACHE[key] = choice
    if cache_file:
        os.makedirs(os.path.dirname(os.path.abspath(cache_file)), exist_ok=True)
        with open(cache_file, "w") as fh:
            json.dump({k: list(v) for k, v in _CACHE.items()}, fh)
    return {"variant": choice, "ms_per_step": {variant_name(k): round(v, 4) for k, v in timings.items()},
            "cached": False}


def autotune_batch(batch, trial_steps: int = 32):
    """Pick the workgroup order (flat or XCD bands) of the batched step for
    every instance of an rfx_amd.batch.Batch by timing ``trial_steps``
    batched steps of each; the batched step runs one cell per thread, so the
    instances are set to the one-cell two-pass variant in that order.
    Leaves the instances reset."""
    n = len(batch.runs)
    step0 = (C.c_int64 * n)(*([0] * n))
    counts = (C.c_int64 * n)(*[min(trial_steps, s) for s in batch.n_steps])
    stream = torch.cuda.current_stream(batch.device)
    timings = {}
    for cand in ((0, 1), (0, 0)):
        for r in batch.runs:
            set_variant(r, *cand)
        for rep in range(2):
            for r in batch.runs:
                r.reset()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            _lib.check(batch.lib.rfxh_batch_run(batch.handle, step0, counts, stream.cuda_stream, 1))
            b.record(stream)
            b.synchronize()
            if rep == 1:
                timings[variant_name(cand)] = a.elapsed_time(b)
    best = min(timings, key=timings.get)
    choice = parse_variant(best)
    for r in batch.runs:
        set_variant(r, *choice)
        r.reset()
    return {"variant": best, "ms": timings}
