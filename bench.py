"""Benchmark: Mcell-steps/s of the FDTD time loop on one MI355X.

Workload (default): 512^3-cell interior, dx = 1 mm, 10-layer CPML on every
face (533^3 = 151,419,437 cells incl. pads), 64 axis-aligned dielectric
boxes at seeded random positions/sizes (8-64 cells per side, eps_r ~ U(1,10),
sigma ~ U(0, 0.05) S/m), an Ez Gaussian-pulse source (f0 = 10 GHz) at the
centre, 16 Ez point probes, one running DFT of Ez on the 533x533 z-plane
through the centre at 32 frequencies, 2000 time steps, float32 fields.

One bench step = one full 2000-time-step run of that problem (fields, CPML
memory and accumulators re-zeroed first); value = steps x 2000 x cells /
wall seconds of the timed region.  Inputs (material ids, coefficient
tables, source table) are resident in HBM before timing starts.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one
GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_STEPS = {"c2": 2000, "c1": 1200}
BYTES_PER_CELL_STEP = 48.0   # 6 float32 field components, one read + one write each
HBM_PEAK_GBS = 8000.0


def c2_spec(n_int=512):
    """The 512^3 workload as high-level inputs (shared by the GPU run, the
    parity tests and the CPU baseline).  n_int < 512 scales the same seeded
    draws (box corners and sizes, probe cells) to an n_int^3 interior."""
    rng = np.random.default_rng(2)
    dx = 1e-3
    L = n_int * dx
    f = n_int / 512
    mats, geom = {}, []
    for b in range(64):
        size = rng.integers(8, 65, size=3)
        lo = np.array([rng.integers(0, 512 - s + 1) for s in size])
        eps = float(rng.uniform(1.0, 10.0))
        sig = float(rng.uniform(0.0, 0.05))
        if n_int != 512:
            size = np.maximum(2, np.round(size * f)).astype(np.int64)
            lo = np.minimum(np.round(lo * f).astype(np.int64), n_int - size)
        name = f"m{b}"
        mats[name] = (eps, sig, 1.0)
        geom.append((tuple((lo + 0.5) * dx), tuple((lo + size + 0.5) * dx), name))
    c = n_int // 2 * dx
    probes = [(tuple(np.round(rng.integers(0, 513, size=3) * f) * dx), "ez") for _ in range(16)]
    return dict(freq_max=20e9, domain=(L, L, L), dx=dx, boundary="cpml", cpml_layers=10,
                materials=mats, geometry=geom,
                sources=[((c, c, c), "ez", 10e9, 0.5)],
                probes=probes,
                dft_planes=[("z", c, "ez", np.linspace(2e9, 20e9, 32), 32)])


def c1_spec():
    """README quick start (87x47x42 cells, 1200 steps)."""
    return dict(freq_max=5e9, domain=(0.14, 0.06, 0.05), dx=2e-3, boundary="cpml",
                cpml_layers=8, materials={"slab": (2.2, 0.01, 1.0)},
                geometry=[((0.07, 0.018, 0.018), (0.09, 0.042, 0.032), "slab")],
                sources=[((0.03, 0.03, 0.025), "ez", 3e9, 0.8)],
                probes=[((0.11, 0.03, 0.025), "ez")], dft_planes=[])


def c4_spec(n=(256, 256, 128), cpml_layers=10):
    """Gradient workload: 256x256x128 interior, dx = 1 mm, 10-layer CPML,
    a line of Ez soft sources along y at (x0, z_c) with a 12 GHz pulse, 3
    Ez probes beyond the design region; objective sum_p sum_f |DFT_f
    ts_p|^2 over 5 frequencies in [11, 13] GHz; 3000 steps.  n scales the
    interior (positions scale with it)."""
    dx = 1e-3
    nx, ny, nz = n
    x0, zc = round(40 * nx / 256) * dx, nz // 2 * dx
    edge = max(2, round(8 * ny / 256))
    sources = [((x0, (j + 0.5) * dx, zc), "ez", 12e9, 0.5) for j in range(edge, ny - edge)]
    off = round(40 * ny / 256)
    probes = [((round(216 * nx / 256) * dx, (ny // 2 + o_) * dx, zc), "ez")
              for o_ in (-off, 0, off)]
    return dict(freq_max=20e9, domain=(nx * dx, ny * dx, nz * dx), dx=dx, boundary="cpml",
                cpml_layers=cpml_layers, materials={}, geometry=[], sources=sources,
                probes=probes, dft_planes=[])


def c4_design(shape, region_n=(64, 64, 4), seed=4):
    """The centred design region (i0, i1, j0, j1, k0, k1) of region_n voxels
    and the eps_r array over the whole grid: 1 outside, U(1, 6) inside."""
    rng = np.random.default_rng(seed)
    eps = np.ones(shape, np.float32)
    c = [s // 2 for s in shape]
    h = [r // 2 for r in region_n]
    region = (c[0] - h[0], c[0] - h[0] + region_n[0], c[1] - h[1], c[1] - h[1] + region_n[1],
              c[2] - h[2], c[2] - h[2] + region_n[2])
    eps[region[0]:region[1], region[2]:region[3], region[4]:region[5]] = \
        rng.uniform(1.0, 6.0, region_n).astype(np.float32)
    return eps, region


C4_FREQS = np.linspace(11e9, 13e9, 5)


def c4_objective_np(ts, dt):
    """sum_p sum_f |sum_n ts[n, p] e^{-j 2 pi f n dt}|^2 in float64."""
    t = np.arange(ts.shape[0]) * dt
    ph = np.exp(-2j * np.pi * C4_FREQS[:, None] * t[None, :])
    acc = ph @ np.asarray(ts, np.float64)
    return float(np.sum(acc.real ** 2 + acc.imag ** 2))


def grad_cost(n_steps=3000):
    """Wall time of forward + reverse sweep over the plain forward run, both
    timed on the device work only (set-up done beforehand)."""
    import torch
    from rfx_amd.adjoint import AdjointRunner
    from rfx_amd.engine import DeviceRun
    spec = c4_spec()
    sim = build_sim(spec)
    grid = sim._build_grid()
    shape = grid.shape
    eps, region = c4_design(shape)
    _, prob, _ = sim._build_problem(n_steps, "forward", eps)
    dt = grid.dt
    freqs = torch.tensor(C4_FREQS, dtype=torch.float64, device="cuda:0")
    t = torch.arange(n_steps, dtype=torch.float64, device="cuda:0") * dt
    ph = torch.exp(-2j * np.pi * freqs[:, None] * t[None, :])      # (5, n)

    def objective(ts):
        acc = ph @ ts.double().to(torch.complex128)                    # (5, probes)
        return (acc.real ** 2 + acc.imag ** 2).sum()

    dr = DeviceRun(prob, n_steps, "cuda:0", two_sets=False)
    dr.run(0, n_steps)
    torch.cuda.synchronize()
    dr.reset()
    t0 = time.perf_counter()
    dr.run(0, n_steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    dr.close()
    # the fastest forward of the same problem: second field set, autotuned
    # (the two-step sweep); the reverse mode records on one set
    from rfx_amd import autotune as at
    db = DeviceRun(prob, n_steps, "cuda:0")
    tune = at.autotune(db, trial_steps=32)
    db.run(0, n_steps)
    torch.cuda.synchronize()
    db.reset()
    t0 = time.perf_counter()
    db.run(0, n_steps)
    torch.cuda.synchronize()
    t_best = time.perf_counter() - t0
    db.close()
    runner = AdjointRunner(prob, n_steps, "cuda:0", region)
    # two gradient passes, the second timed: the first allocates the
    # reverse-mode buffers and uploads the derivative tables, as the
    # forward's untimed first run does its set-up
    for _ in range(2):
        runner.dr.reset()
        runner.checkpoints = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ts = runner.forward().detach().requires_grad_(True)
        J = objective(ts)
        (g,) = torch.autograd.grad(J, ts)
        grad = runner.backward(g)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
    finite = bool(torch.isfinite(grad).all())
    del runner
    torch.cuda.empty_cache()
    return {"workload": f"{shape[0]}x{shape[1]}x{shape[2]} cells, 64x64x4 design region, "
                        f"{len(spec['sources'])} line-source cells, 3 probes, 5-freq DFT "
                        f"objective, {n_steps} time steps, one record segment, second of two "
                        f"gradient passes",
            "forward_s": round(t_fwd, 4), "forward_adjoint_s": round(t_grad, 4),
            "ratio": round(t_grad / t_fwd, 3), "gradient_finite": finite,
            "forward_variant": "single field set, default two-pass step",
            "fastest_forward_s": round(t_best, 4),
            "fastest_forward_variant": at.variant_name(tuple(tune["variant"])),
            "ratio_vs_fastest_forward": round(t_grad / t_best, 3)}


def c3_profile():
    """dz (m) of the graded patch workload, 120 cells: 10 x 1.0 mm air,
    13 cells geometric 1.0 -> 0.1 mm (ratio 1.194), 16 x 0.1 mm substrate,
    13 cells geometric 0.1 -> 1.0 mm, 68 x 1.0 mm air."""
    r = 0.1 ** (1.0 / 13.0)
    down = [1e-3 * r ** k for k in range(1, 14)]
    up = [1e-4 / r ** k for k in range(1, 14)]
    dz = [1e-3] * 10 + down + [1e-4] * 16 + up + [1e-3] * 68
    return np.asarray(dz)


def c3_spec(n_xy=300):
    """Graded patch antenna: dx = dy = 0.5 mm, n_xy x n_xy interior, graded
    z (c3_profile), 12-layer CPML; FR4-like substrate (eps_r 4.4, sigma
    0.002) between a PEC ground sheet and a 30 x 38 mm PEC patch sheet
    (sigma >= 1e6 -> mask); a 50-ohm wire port from ground to patch
    (1.6 mm), 1-4 GHz pulse (f0 2.5 GHz, bandwidth 1.2); S11 at 201
    frequencies; 8000 steps."""
    dx = 5e-4
    dz = c3_profile()
    L = n_xy * dx
    z0 = float(np.sum(dz[:23]))           # substrate bottom node
    z1 = z0 + 16 * 1e-4                   # substrate top node
    c = L / 2
    return dict(freq_max=4e9, domain=(L, L), dx=dx, dz_profile=dz, cpml_layers=12,
                materials={"fr4": (4.4, 0.002, 1.0), "pec": (1.0, 1e10, 1.0)},
                geometry=[((0.0, 0.0, z0), (L, L, z1), "fr4"),
                          ((0.0, 0.0, z0), (L, L, z0), "pec"),
                          ((c - 0.015, c - 0.019, z1), (c + 0.015, c + 0.019, z1), "pec")],
                port=((c - 0.005, c, z0), 50.0, 1.6e-3, 2.5e9, 1.2),
                s_freqs=np.linspace(1e9, 4e9, 201), probes=[((c + 0.02, c, z1 + 2e-3), "ez")])


def build_c3_sim(spec):
    import rfx_amd as R
    sim = R.Simulation(spec["freq_max"], spec["domain"], dx=spec["dx"], boundary="cpml",
                       cpml_layers=spec["cpml_layers"], dz_profile=spec["dz_profile"])
    for name, (e, s_, m) in spec["materials"].items():
        sim.add_material(name, eps_r=e, sigma=s_, mu_r=m)
    for lo, hi, name in spec["geometry"]:
        sim.add(R.Box(lo, hi), material=name)
    pos, z0, ext, f0, bw = spec["port"]
    sim.add_port(pos, "ez", impedance=z0, extent=ext,
                 waveform=R.GaussianPulse(f0=f0, bandwidth=bw))
    for pos, comp in spec["probes"]:
        sim.add_probe(pos, comp)
    return sim


def graded_run(n_steps=8000):
    """Throughput of the graded lane on the patch workload (whole run incl.
    wire-port sampling; S11 from the port DFT)."""
    import torch
    from rfx_amd.engine import DeviceRun
    spec = c3_spec()
    sim = build_c3_sim(spec)
    grid, prob, ports = sim._build_problem(n_steps, "run",
                                           s11_freqs=np.asarray(spec["s_freqs"], np.float32))
    dr = DeviceRun(prob, n_steps, "cuda:0", two_sets=False)
    from rfx_amd import autotune as at
    tune = at.autotune(dr, trial_steps=32)
    dr.run(0, min(n_steps, 64))
    torch.cuda.synchronize()
    dr.reset()
    t0 = time.perf_counter()
    dr.run(0, n_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    cells = int(np.prod(grid.shape))
    finite = bool(torch.isfinite(dr.port_samples()).all())
    dr.close()
    return {"workload": f"graded patch {grid.shape[0]}x{grid.shape[1]}x{grid.shape[2]} cells, "
                        f"dz {float(grid.d64[2].min()) * 1e3:.2f}-{float(grid.d64[2].max()) * 1e3:.2f}"
                        f" mm, wire port, {n_steps} steps",
            "variant": at.variant_name(tuple(tune["variant"])),
            "Mcell_steps_per_s": round(n_steps * cells / el / 1e6, 1), "seconds": round(el, 3),
            "port_samples_finite": finite}


def c5_specs(n_inst=48, sizes=None, steps=(1500, 4001), thick=(4, 21), seed=5):
    """Batched-sweep workload: 48 instances, interior sizes from {64^3, 96^3,
    128x128x64, 160x96x96, 192x192x48}, dx = 1 mm, 8-layer CPML, one
    dielectric slab (eps_r ~ U(2,12), 4-20 cells thick, normal to a random
    axis), an Ez pulse source (f0 ~ U(2,20) GHz), 4 Ez probes, 1500-4000
    steps; seed 5.  (sizes, steps, thick, seed: the same recipe at other
    sizes, for tests.)"""
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = [(64, 64, 64), (96, 96, 96), (128, 128, 64), (160, 96, 96), (192, 192, 48)]
    dx = 1e-3
    out = []
    for q in range(n_inst):
        n = sizes[int(rng.integers(0, len(sizes)))]
        ax = int(rng.integers(0, 3))
        th = int(rng.integers(*thick))
        c0 = int(rng.integers(1, n[ax] - th - 1))
        lo, hi = [0.0, 0.0, 0.0], [n[0] * dx, n[1] * dx, n[2] * dx]
        lo[ax], hi[ax] = (c0 + 0.5) * dx, (c0 + th + 0.5) * dx
        eps = float(rng.uniform(2, 12))
        f0 = float(rng.uniform(2e9, 20e9))
        n_steps = int(rng.integers(*steps))
        src = tuple(0.3 * v * dx for v in n)
        probes = [(tuple(float(rng.uniform(0.2, 0.8)) * v * dx for v in n), "ez") for _ in range(4)]
        out.append((dict(freq_max=25e9, domain=tuple(v * dx for v in n), dx=dx, boundary="cpml",
                         cpml_layers=8, materials={"slab": (eps, 0.0, 1.0)},
                         geometry=[(tuple(lo), tuple(hi), "slab")],
                         sources=[(src, "ez", f0, 0.5)], probes=probes, dft_planes=[]), n_steps))
    return out


def batch_sweep():
    """The batched-sweep workload: 48 forwards, and for instances 0, 2, 4, ...
    d(sum of squared probe samples)/d eps_r over the whole grid.  Timed
    batched (one launch per sub-step over all active instances; the
    gradient instances by batched reverse mode) and one instance after
    another (each with its own captured graphs; per-instance reverse mode)."""
    import torch
    from rfx_amd.batch import Batch
    from rfx_amd.adjoint import AdjointRunner, batch_gradients
    specs = c5_specs()
    sims = [build_sim(sp) for sp, _ in specs]
    steps = [n for _, n in specs]
    b = Batch(sims, steps, "cuda:0")
    cell_steps = sum(int(np.prod(r.prob.shape)) * n for r, n in zip(b.runs, steps))
    from rfx_amd import autotune as at
    tune = at.autotune_batch(b, trial_steps=64)
    b.run()                                   # warm-up: graphs captured
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    for r, n in zip(b.runs, steps):           # warm-up of the per-instance graphs
        at.set_variant(r, 4, 1)               # each instance on its own default step
        r.reset()
        r.run(0, min(n, 64))
    torch.cuda.synchronize()
    t_one = []
    for r, n in zip(b.runs, steps):
        t0 = time.perf_counter()
        r.reset()
        r.run(0, n)
        torch.cuda.synchronize()
        t_one.append(time.perf_counter() - t0)
    t_seq = sum(t_one)
    problems = [r.prob for r in b.runs]
    b.close()
    del b
    torch.cuda.empty_cache()
    grad_ids = list(range(0, len(problems), 2))
    other = [q for q in range(len(problems)) if q not in grad_ids]
    # the forward-only half, batched
    fb = Batch([sims[q] for q in other], [steps[q] for q in other], "cuda:0")
    at.autotune_batch(fb, trial_steps=64)
    fb.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fb.run()
    torch.cuda.synchronize()
    t_other = time.perf_counter() - t0
    fb.close()
    del fb
    torch.cuda.empty_cache()
    # the gradient half, batched reverse mode (forward with checkpoints,
    # recomputation with recording, reverse sweep; setup included)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ts_b, grads_b = batch_gradients([problems[q] for q in grad_ids], [steps[q] for q in grad_ids],
                                    lambda q, ts: 2.0 * ts, "cuda:0")
    torch.cuda.synchronize()
    t_adj_b = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(g).all()) for g in grads_b)
    del ts_b, grads_b
    torch.cuda.empty_cache()
    # the gradient half, one instance after another
    t0 = time.perf_counter()
    for q in grad_ids:
        runner = AdjointRunner(problems[q], steps[q], "cuda:0")
        ts = runner.forward()
        grad = runner.backward(2.0 * ts)
        runner.dr.close()
        del runner, grad
    torch.cuda.synchronize()
    t_adj_s = time.perf_counter() - t0
    torch.cuda.empty_cache()
    whole_b = t_other + t_adj_b
    whole_s = sum(t_one[q] for q in other) + t_adj_s
    return {"workload": f"{len(specs)} instances, {min(steps)}-{max(steps)} steps, "
                        f"{cell_steps / 1e9:.1f} G cell-steps; d(sum ts^2)/d eps_r of "
                        f"{len(grad_ids)} of them",
            "batched_variant": tune["variant"],
            "batched_s": round(t_batch, 3), "sequential_s": round(t_seq, 3),
            "batched_Mcell_steps_per_s": round(cell_steps / t_batch / 1e6, 1),
            "sequential_Mcell_steps_per_s": round(cell_steps / t_seq / 1e6, 1),
            "adjoint_instances": len(grad_ids),
            "adjoint_batched_s": round(t_adj_b, 3),
            "adjoint_sequential_s": round(t_adj_s, 3),
            "adjoint_gain": round(t_adj_s / t_adj_b, 2),
            "gradients_finite": finite,
            "whole_workload_batched_s": round(whole_b, 3),
            "whole_workload_sequential_s": round(whole_s, 3),
            "whole_workload_gain": round(whole_s / whole_b, 2)}


def build_sim(spec):
    import rfx_amd as R
    sim = R.Simulation(spec["freq_max"], spec["domain"], dx=spec["dx"],
                       boundary=spec["boundary"], cpml_layers=spec["cpml_layers"])
    for name, (e, s, m) in spec["materials"].items():
        sim.add_material(name, eps_r=e, sigma=s, mu_r=m)
    for lo, hi, name in spec["geometry"]:
        sim.add(R.Box(lo, hi), material=name)
    for pos, comp, f0, bw in spec["sources"]:
        sim.add_source(pos, comp, waveform=R.GaussianPulse(f0=f0, bandwidth=bw))
    for pos, comp in spec["probes"]:
        sim.add_probe(pos, comp)
    for axis, coord, comp, freqs, nf in spec["dft_planes"]:
        sim.add_dft_plane_probe(axis=axis, coordinate=coord, component=comp, freqs=freqs,
                                n_freqs=nf)
    return sim


def oracle_spec(spec):
    from oracle import fdtd_np as o
    return dict(freq_max=spec["freq_max"], domain=spec["domain"], dx=spec["dx"],
                boundary=spec["boundary"], cpml_layers=spec["cpml_layers"],
                materials=spec["materials"], geometry=spec["geometry"],
                ports=[(pos, comp, 0.0, o.Pulse(f0, bw), True)
                       for pos, comp, f0, bw in spec["sources"]],
                probes=spec["probes"], dft_planes=spec["dft_planes"], lane="run")


_CPU_CHILD = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1])
import numpy as np
import bench
from oracle import cstep
from oracle import fdtd_np as o
n1, n2 = int(sys.argv[2]), int(sys.argv[3])
spec = bench.c2_spec() if sys.argv[4] == "c2" else bench.c1_spec()
out = {}
for n in (n1, n2):
    g, cfg = o.build_uniform(bench.oracle_spec(spec), n)
    t0 = time.perf_counter()
    cstep.run_c(cfg, n, double=False, want_fields=False)
    out[n] = time.perf_counter() - t0
print(json.dumps({"cells": int(np.prod(g.shape)), "shape": list(g.shape),
                  "t": [out[n1], out[n2]]}))
"""


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rate(config, threads, n1, n2):
    """Mcell-steps/s of the C restatement (oracle/cstep.c, float32, OpenMP)
    with `threads` threads: two runs of the same workload, n1 and n2 time
    steps, in a child process; the rate is (n2 - n1) steps over the
    difference of their times, so allocation and first-touch set-up cancel."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    r = subprocess.run([sys.executable, "-c", _CPU_CHILD, ROOT, str(n1), str(n2), config],
                       env=env, capture_output=True, text=True, check=True)
    d = json.loads(r.stdout.strip().splitlines()[-1])
    dt = d["t"][1] - d["t"][0]
    return (n2 - n1) * d["cells"] / dt / 1e6, d, dt


def cpu_baseline(config, n_total, all_steps, one_steps):
    """The C restatement timed on this host's cores, on rank 0: all the
    cores this process may use, and one core."""
    cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    v_all, d, t_all = cpu_rate(config, cores, 1, 1 + all_steps)
    v_one, _, t_one = cpu_rate(config, 1, 1, 1 + one_steps)
    shape = "x".join(str(v) for v in d["shape"])
    return {"value": round(v_all, 2), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "one_core": {"value": round(v_one, 2), "unit": "Mcell-steps/s", "cores": 1,
                         "timed_steps": one_steps, "seconds": round(t_one, 2)},
            "sample": f"{all_steps} time steps ({cores} OpenMP threads, {t_all:.1f} s) and "
                      f"{one_steps} (1 thread, {t_one:.1f} s) of the same {shape} problem "
                      f"of {n_total} steps, C restatement (oracle/cstep.c) in float32: each "
                      f"the difference of a run of 1 + k steps and a run of 1 step, so set-up "
                      f"and first touch cancel"}


def c1_line():
    """The README quick start (87x47x42 cells, 1200 steps) on the product
    path: autotuned variant, graph replay, inputs resident."""
    import torch
    from rfx_amd import autotune as at
    from rfx_amd.engine import DeviceRun
    spec = c1_spec()
    n = N_STEPS["c1"]
    sim = build_sim(spec)
    grid, prob, _ = sim._build_problem(n, "run")
    dr = DeviceRun(prob, n, "cuda:0")
    tune = at.autotune(dr, trial_steps=64)
    for _ in range(2):
        dr.reset()
        dr.run(0, n)
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        dr.reset()
        dr.run(0, n)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    finite = bool(torch.isfinite(dr.time_series()).all())
    dr.close()
    cells = int(np.prod(grid.shape))
    return {"workload": f"README quick start {grid.shape[0]}x{grid.shape[1]}x{grid.shape[2]} "
                        f"cells, {n} time steps, graph replay, mean of {reps} runs",
            "variant": at.variant_name(tuple(tune["variant"])),
            "Mcell_steps_per_s": round(n * cells / el / 1e6, 1),
            "ms_per_run": round(el * 1e3, 3), "us_per_time_step": round(el / n * 1e6, 2),
            "series_finite": finite}


def kernel_roofline(kern_v, ms_h, ms_e, cells):
    """The dominant kernel of the variant, its average launch time and its
    algorithmic bytes per launch (48 B per cell-step)."""
    if kern_v == 8:
        # one launch advances two steps: 2 x 48 B per cell
        return "k_step_tb2", ms_h, 2 * BYTES_PER_CELL_STEP * cells
    if kern_v in (1, 2, 3, 7):
        # one launch does the whole step: 48 B per cell
        which = {1: "k_step_fused", 2: "k_step_march", 3: "k_step_march",
                 7: "k_step_wmarch"}[kern_v]
        return which, ms_h, BYTES_PER_CELL_STEP * cells
    # two passes, each advancing half the field components: 24 B per cell
    sfx = "_col" if kern_v in (4, 5, 6) else ""
    which, ms = (f"k_update_e{sfx}", ms_e) if ms_e >= ms_h else (f"k_update_h{sfx}", ms_h)
    return which, ms, 0.5 * BYTES_PER_CELL_STEP * cells


def traffic_of(config, variant, which):
    """HBM bytes per launch of `which` from the committed counter profile of
    this workload and variant (the profile named beside the figure), or None."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as fh:
        tab = json.load(fh)
    rec = tab.get(f"{config}:{variant}:{which}")
    if rec is None and which == "k_step_tb2":
        rec = tab.get(f"{config}:tb2:{which}")
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(N_STEPS))
    ap.add_argument("--cpu-steps", type=int, default=12,
                    help="time steps of the all-core CPU sample (the 1-core sample takes "
                         "a twelfth of them, at least one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graded", action="store_true",
                    help="skip the graded-lane timing")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-sweep timing")
    ap.add_argument("--no-grad", action="store_true",
                    help="skip the forward+adjoint / forward timing")
    ap.add_argument("--no-c1", action="store_true", help="skip the README quick-start line")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-kernel-sample", action="store_true",
                    help="skip the event-sampled run that times the kernels")
    ap.add_argument("--time-steps", type=int, default=None,
                    help="time steps per run (default: the workload's own count); "
                         "for short profiled runs")
    ap.add_argument("--variant", default="auto",
                    help="kernel variant; auto = the project autotuner's choice for this "
                         "problem (timed on its first steps, state reset afterwards)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 512 MiB device copy five times, the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from rfx_amd.engine import DeviceRun
    spec = c2_spec() if args.config == "c2" else c1_spec()
    n_total = args.time_steps or N_STEPS[args.config]
    t_setup = time.perf_counter()
    sim = build_sim(spec)
    grid, prob, _ = sim._build_problem(n_total, "run")
    dr = DeviceRun(prob, n_total, "cuda:0")
    torch.cuda.synchronize()
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f} s, grid {grid.shape}, "
          f"materials {len(prob.mat.table)}", file=sys.stderr, flush=True)
    from rfx_amd import autotune as at
    tuning = None
    if args.variant == "auto":
        tuning = at.autotune(dr, trial_steps=32)
        kern_v, prm_v = tuning["variant"]
    else:
        kern_v, prm_v = at.parse_variant(args.variant)
        at.set_variant(dr, kern_v, prm_v)
    variant = at.variant_name((kern_v, prm_v))
    print(f"[bench] variant {variant} (autotune: {tuning})", file=sys.stderr, flush=True)
    cells = int(np.prod(grid.shape))
    if args.copy_calib:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print("[bench] copy calibration: 5 x 512 MiB device copies", file=sys.stderr,
              flush=True)
    use_graph = not args.no_graph
    for w in range(args.warmup):
        dr.reset()
        dr.run(0, n_total, use_graph=use_graph)
        torch.cuda.synchronize()
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    # the timed region: the product path (hipGraph replay of the step
    # sequence), no event sampling
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        dr.reset()
        dr.run(0, n_total, use_graph=use_graph)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} runs in {el:.3f} s", file=sys.stderr, flush=True)
    finite = bool(torch.isfinite(dr.time_series()).all()) and all(
        bool(torch.isfinite(a).all()) for a in dr.plane_acc)
    # kernel durations: HIP events on the run's stream around up to 64
    # evenly spaced launches of one more run (direct launches: events
    # cannot sit inside a replayed graph)
    ms_h = ms_e = ms_d = 0.0
    n_e = 0
    if not args.no_kernel_sample:
        dr.reset()
        dr.set_timing(True)
        dr.run(0, n_total, use_graph=False)
        torch.cuda.synchronize()
        ms_h, n_h = dr.kernel_time_ms(0)
        ms_e, n_e = dr.kernel_time_ms(1)
        ms_d, n_d = dr.kernel_time_ms(2)
        dr.set_timing(False)      # (clears the sums)
    value = args.steps * n_total * cells / el / 1e6
    which, ms, unit_bytes = kernel_roofline(kern_v, ms_h, ms_e, cells)
    achieved = unit_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None
    trec = traffic_of(args.config, variant, which)
    launches_per_step = 0.5 if kern_v == 8 else 1.0
    out = {
        "metric": "Mcell-steps/s fp32 + % HBM roofline; fwd+adjoint grad time vs plain forward",
        "value": round(value, 1), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"uniform Yee + 10-layer CPML, {grid.shape[0]}x{grid.shape[1]}x"
                                f"{grid.shape[2]} cells, 64 random dielectric boxes, Ez source, "
                                f"16 probes, 32-freq DFT plane, {n_total} time steps")
                   if args.config == "c2" else
                   f"README quick start {grid.shape}, {n_total} time steps",
                   "cells": cells, "time_steps": n_total, "variant": variant,
                   "timed_path": "hipGraph replay" if use_graph else "direct launches",
                   "ms_per_time_step": round(el / args.steps / n_total * 1e3, 4),
                   "outputs_finite": finite,
                   "autotune_ms_per_step": tuning["ms_per_step"] if tuning else None,
                   "pct_hbm_roofline_48B": round(100 * value * 1e6 * BYTES_PER_CELL_STEP
                                                 / (HBM_PEAK_GBS * 1e9), 2),
                   "kernel_ms_event_sampled": {
                       "step_kernel_or_h_pass": round(ms_h, 4), "e_pass": round(ms_e, 4),
                       "dft_and_counter": round(ms_d, 4), "samples": n_e,
                       "time_steps_per_launch": 1 / launches_per_step}},
        "roofline": {"bound": "hbm", "kernel": which,
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "algorithmic_bytes_per_launch": unit_bytes,
                     "traffic": trec["bytes_per_launch"] if trec else None,
                     "traffic_profile": trec["profile"] if trec else None},
        "cpu_baseline": None,
        "grad_vs_forward": None,
        "batch_sweep": None,
        "graded_lane": None,
        "readme_quick_start": None,
    }
    dr.close()
    del dr
    torch.cuda.empty_cache()
    if not args.no_c1 and args.config == "c2":
        print("[bench] README quick start ...", file=sys.stderr, flush=True)
        out["readme_quick_start"] = c1_line()
        print(f"[bench] c1 {out['readme_quick_start']}", file=sys.stderr, flush=True)
    if not args.no_grad and args.config == "c2":
        print("[bench] forward+adjoint vs forward ...", file=sys.stderr, flush=True)
        out["grad_vs_forward"] = grad_cost()
        print(f"[bench] grad {out['grad_vs_forward']}", file=sys.stderr, flush=True)
    if not args.no_graded and args.config == "c2":
        print("[bench] graded lane ...", file=sys.stderr, flush=True)
        out["graded_lane"] = graded_run()
        print(f"[bench] graded {out['graded_lane']}", file=sys.stderr, flush=True)
    if not args.no_batch and args.config == "c2":
        print("[bench] batched sweep ...", file=sys.stderr, flush=True)
        out["batch_sweep"] = batch_sweep()
        print(f"[bench] batch {out['batch_sweep']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.config, n_total, args.cpu_steps,
                                           max(1, args.cpu_steps // 12))
        print(f"[bench] cpu {out['cpu_baseline']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
