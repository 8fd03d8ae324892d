"""Shared problem descriptions for the oracle and parity tests.

Each case is described once, as the high-level inputs a user gives the
``Simulation`` API, and is turned into an oracle spec here and into
``rfx_amd.Simulation`` calls in the GPU tests.
"""
import json
import os

import numpy as np

from oracle import fdtd_np as o

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# rfx tests/test_lumped_rlc_ad.py:135-151
S11_CASE = dict(freq_max=10e9, domain=(0.02, 0.02, 0.02), dx=0.02 / 15,
                boundary="cpml", cpml_layers=6,
                ports=[((0.0093,) * 3, "ez", 50.0, o.Pulse(5e9, 0.9), True)],
                lane="forward", s11_freqs=np.array([3., 4., 5., 6., 7.]) * 1e9)
S11_STEPS = 1200


def golden_s11():
    return np.load(os.path.join(GOLDEN, "golden_forward_no_rlc_s11.npy"))


def ladder():
    with open(os.path.join(GOLDEN, "ad_fd_ladder_values.json")) as fh:
        return json.load(fh)


# scripts/diagnostics/gradient_dx_ladder/run_ladder.py:60-165
LADDER_LAMBDA = 2.998e8 / 10e9


def ladder_case(n_per_lambda, eps_val):
    dx = LADDER_LAMBDA / n_per_lambda
    g = o.UniformGrid(10e9, (0.06, 0.012, 0.012), dx, 8)
    n_steps = g.num_timesteps(20.0)
    xs = (np.arange(g.shape[0]) - g.pad) * dx
    m1d = (xs >= 30e-3) & (xs < 36e-3)
    eps = np.where(m1d[:, None, None], np.float32(eps_val), np.float32(1.0)) * np.ones(g.shape, np.float32)
    spec = dict(freq_max=10e9, domain=(0.06, 0.012, 0.012), dx=dx, boundary="cpml",
                cpml_layers=8,
                ports=[((12e-3, 6e-3, 6e-3), "ez", 0.0, o.Pulse(8e9, 0.9), True)],
                probes=[((48e-3, 6e-3, 6e-3), "ez")], lane="forward", eps_override=eps)
    return spec, n_steps, m1d


def ladder_objective(ts, dt, f_obj=8e9):
    t = np.arange(ts.shape[0]) * dt
    acc = np.sum(ts * np.exp(-1j * 2 * np.pi * f_obj * t))
    return float(abs(acc) ** 2)


# README quick start (rfx README.md:49-74)
def quickstart_case():
    return dict(freq_max=5e9, domain=(0.14, 0.06, 0.05), dx=2e-3, boundary="cpml",
                cpml_layers=8, materials={"slab": (2.2, 0.01, 1.0)},
                geometry=[((0.07, 0.018, 0.018), (0.09, 0.042, 0.032), "slab")],
                ports=[((0.03, 0.03, 0.025), "ez", 0.0, o.Pulse(3e9, 0.8), True)],
                probes=[((0.11, 0.03, 0.025), "ez")], lane="run")


# a small case touching every feature of the uniform lane
def mixed_case(lane="run"):
    return dict(freq_max=10e9, domain=(0.016, 0.012, 0.010), dx=1e-3, boundary="cpml",
                cpml_layers=4,
                materials={"d": (3.0, 0.02, 1.0), "m": (1.0, 1e8, 1.0), "mag": (1.5, 0.0, 2.0)},
                geometry=[((0.004, 0.003, 0.002), (0.009, 0.008, 0.006), "d"),
                          ((0.011, 0.002, 0.003), (0.0125, 0.009, 0.0035), "m"),
                          ((0.002, 0.008, 0.006), (0.005, 0.0105, 0.009), "mag")],
                ports=[((0.003, 0.006, 0.005), "ez", 0.0, o.Pulse(8e9, 0.8), True),
                       ((0.0125, 0.006, 0.005), "ey", 0.0, o.Pulse(6e9, 0.5), True),
                       ((0.008, 0.004, 0.003), "ex", 50.0, o.Pulse(7e9, 0.9), True)],
                probes=[((0.013, 0.006, 0.005), "ez"), ((0.010, 0.005, 0.004), "hx"),
                        ((0.003, 0.006, 0.005), "ez"), ((0.006, 0.007, 0.002), "hz")],
                dft_planes=[("z", 0.005, "ez", None, 4), ("x", 0.008, "hy", None, 3)],
                s11_freqs=np.array([5e9, 7e9, 9e9]), lane=lane)


def to_sim(spec, device=None):
    """Build an rfx_amd.Simulation from an oracle spec (same registration order)."""
    import rfx_amd as R
    kw = {}
    if device is not None:
        kw["device"] = device
    for key in ("dx_profile", "dy_profile", "dz_profile"):
        if spec.get(key) is not None:
            kw[key] = spec[key]
    sim = R.Simulation(spec["freq_max"], spec["domain"], dx=spec.get("dx"),
                       boundary=spec.get("boundary", "cpml"),
                       cpml_layers=spec.get("cpml_layers", 16),
                       cpml_kappa_max=spec.get("kappa_max", 1.0), **kw)
    for name, (e, s, m) in spec.get("materials", {}).items():
        poles = [R.DebyePole(de, tau) for de, tau in spec.get("debye", {}).get(name, [])]
        lor = [R.LorentzPole(*p) for p in spec.get("lorentz", {}).get(name, [])]
        sim.add_material(name, eps_r=e, sigma=s, mu_r=m, debye_poles=poles or None,
                         lorentz_poles=lor or None)
    for lo, hi, name in spec.get("geometry", []):
        sim.add(R.Box(lo, hi), material=name)
    for entry in spec.get("ports", []):
        pos, comp, z0, pulse, excite = entry[:5]
        extent = entry[5] if len(entry) > 5 else None
        wf = None if pulse is None else R.GaussianPulse(pulse.f0, pulse.bandwidth,
                                                         pulse.amplitude, pulse.cutoff)
        if z0 == 0.0:
            sim.add_source(pos, comp, waveform=wf)
        else:
            sim.add_port(pos, comp, impedance=z0, waveform=wf, excite=excite, extent=extent)
    for pos, comp in spec.get("probes", []):
        sim.add_probe(pos, comp)
    for axis, coord, comp, freqs, nf in spec.get("dft_planes", []):
        sim.add_dft_plane_probe(axis=axis, coordinate=coord, component=comp, freqs=freqs,
                                n_freqs=nf)
    for axis, coord, freqs, nf, size, center, window, alpha in spec.get("flux_monitors", []):
        sim.add_flux_monitor(axis=axis, coordinate=coord, freqs=freqs, n_freqs=nf, size=size,
                             center=center, dft_window=window, dft_window_alpha=alpha)
    if spec.get("ntff") is not None:
        lo, hi, freqs, nf = spec["ntff"]
        sim.add_ntff_box(lo, hi, freqs=freqs, n_freqs=nf)
    return sim


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def smooth_grading(cells, max_ratio=1.3):
    """Geometric transition cells wherever adjacent cells differ by more
    than max_ratio (rfx/auto_config.py:656-723, no preserved regions)."""
    cells = [float(c) for c in cells]
    out = [cells[0]]
    for target in cells[1:]:
        prev = out[-1]
        while target / prev > max_ratio + 1e-12:
            prev = prev * max_ratio
            out.append(prev)
        while prev / target > max_ratio + 1e-12:
            prev = prev / max_ratio
            out.append(prev)
        out.append(target)
    return np.array(out)


# Air-filled PEC cavity on a 4:1 graded z mesh
# (tests/test_nonuniform_cavity_accuracy.py:80-116 of the reference)
CAVITY_A, CAVITY_B, CAVITY_DX = 40e-3, 35e-3, 1e-3


def graded_cavity_case(n_steps=8000):
    fine = 0.25e-3
    dz_raw = [CAVITY_DX] * 17 + [fine] * 8 + [CAVITY_DX] * 17
    dz = smooth_grading(dz_raw, 1.3)
    d = float(np.sum(dz))
    f111 = (o.C0 / 2) * np.sqrt((1 / CAVITY_A) ** 2 + (1 / CAVITY_B) ** 2 + (1 / d) ** 2)
    a, b = CAVITY_A, CAVITY_B
    spec = dict(freq_max=2 * f111, domain=(a, b), dx=CAVITY_DX, dz_profile=dz,
                boundary="pec",
                ports=[((a / 3, b / 3, d / 4), "ez", 0.0, o.Pulse(f111, 0.8), True)],
                probes=[((2 * a / 3, 2 * b / 3, d / 4), "ez")])
    return spec, n_steps, f111, d


def mixed_graded_case():
    """Graded-z case touching every feature of the graded lane: CPML on all
    faces, lossy / magnetic boxes, a PEC sheet, a soft source, a single-cell
    port, a wire port with V/I, E and H probes, DFT planes."""
    dz = smooth_grading([1e-3] * 4 + [0.4e-3] * 4 + [1e-3] * 4, 1.3)
    d = float(np.sum(dz))
    return dict(freq_max=10e9, domain=(0.016, 0.012), dx=1e-3, dz_profile=dz,
                boundary="cpml", cpml_layers=4,
                materials={"sub": (4.4, 0.02, 1.0), "pec": (1.0, 1e10, 1.0),
                           "mag": (1.5, 0.0, 2.0)},
                geometry=[((0.003, 0.002, 0.0), (0.013, 0.010, 0.004), "sub"),
                          ((0.004, 0.003, 0.004), (0.009, 0.007, 0.004), "pec"),
                          ((0.011, 0.008, 0.007), (0.014, 0.011, 0.010), "mag")],
                ports=[((0.006, 0.005, 0.002), "ez", 0.0, o.Pulse(8e9, 0.8), True),
                       ((0.012, 0.004, 0.003), "ez", 50.0, o.Pulse(6e9, 0.7), True, 0.0022),
                       ((0.003, 0.009, 0.008), "ex", 50.0, None, False)],
                probes=[((0.010, 0.006, 0.004), "ez"), ((0.005, 0.008, 0.006), "hy"),
                        ((0.012, 0.004, 0.003), "ez")],
                dft_planes=[("z", 0.004, "ez", None, 3), ("x", 0.008, "hx", None, 2)],
                s11_freqs=np.array([5e9, 7e9]))


def sweep_instance(rng, shape_cells, dx=1e-3, layers=4):
    """One instance of a batched sweep: vacuum box with a random dielectric
    slab normal to a random axis, an Ez pulse source and four Ez probes
    (the sweep workload, at test sizes)."""
    nx, ny, nz = shape_cells
    dom = (nx * dx, ny * dx, nz * dx)
    ax = int(rng.integers(0, 3))
    th = int(rng.integers(2, 6))
    lo = [0.0, 0.0, 0.0]
    hi = list(dom)
    c0 = int(rng.integers(2, shape_cells[ax] - th - 1))
    lo[ax] = (c0 + 0.5) * dx
    hi[ax] = (c0 + th + 0.5) * dx
    f0 = float(rng.uniform(2e9, 20e9))
    src = tuple(0.3 * d for d in dom)
    probes = [(tuple(float(rng.uniform(0.2, 0.8)) * d for d in dom), "ez") for _ in range(4)]
    return dict(freq_max=25e9, domain=dom, dx=dx, boundary="cpml", cpml_layers=layers,
                materials={"slab": (float(rng.uniform(2, 12)), 0.0, 1.0)},
                geometry=[(tuple(lo), tuple(hi), "slab")],
                ports=[(src, "ez", 0.0, o.Pulse(f0, 0.5), True)],
                probes=probes, lane="run")


def flux_box(h, c=0.015, freqs=(8e9, 10e9, 12e9)):
    """Six finite flux monitors forming a closed box of half-size h around
    (c, c, c): (axis, coord, freqs, n, size, center, window, alpha)."""
    out = []
    for name in "xyz":
        for sgn in (-1, 1):
            out.append((name, c + sgn * h, np.array(freqs, np.float32), len(freqs),
                        (2 * h, 2 * h), (c, c), "rect", 0.25))
    return out


def flux_case():
    return dict(freq_max=20e9, domain=(0.03, 0.03, 0.03), dx=1e-3, boundary="cpml",
                cpml_layers=8, ports=[((0.015, 0.015, 0.015), "ez", 0.0, o.Pulse(10e9, 0.5), True)],
                probes=[((0.02, 0.015, 0.015), "ez")], lane="run",
                flux_monitors=flux_box(0.005) + flux_box(0.009))


def ntff_case():
    """flux_case with an NTFF box of half-size 7 mm between its two flux
    boxes, at the flux monitors' frequencies."""
    spec = flux_case()
    spec["ntff"] = ((0.008, 0.008, 0.008), (0.022, 0.022, 0.022),
                    np.array((8e9, 10e9, 12e9), np.float32), 3)
    return spec


def dipole_ntff_case():
    """The reference's short-dipole NTFF set-up (tests/test_farfield.py:
    24-56 of the reference): 30 mm cube, 8-layer CPML, default dx at
    5 GHz, Ez pulse at 3 GHz in the centre, NTFF box 4 mm inside the domain
    edges, one frequency; 200 steps."""
    return dict(freq_max=5e9, domain=(0.03, 0.03, 0.03), boundary="cpml", cpml_layers=8,
                ports=[((0.015, 0.015, 0.015), "ez", 0.0, o.Pulse(3e9, 0.5), True)],
                probes=[((0.015, 0.015, 0.015), "ez")], lane="run",
                ntff=((0.004, 0.004, 0.004), (0.026, 0.026, 0.026),
                      np.array([3e9], np.float32), 1)), 200


def debye_case(tau, delta_eps=3.0, eps_inf=2.0, lane="run", overlap=True):
    """Small CPML case with a Debye box (eps_inf, one pole delta_eps, tau),
    optionally a second box of another material overlapping it (the pole
    stays where the later box sets the permittivity), a lumped port and a
    soft source outside the Debye box, and E / H probes."""
    geometry = [((0.004, 0.003, 0.002), (0.011, 0.009, 0.007), "deb")]
    if overlap:
        geometry.append(((0.009, 0.002, 0.001), (0.013, 0.006, 0.005), "d"))
    return dict(freq_max=10e9, domain=(0.016, 0.012, 0.010), dx=1e-3, boundary="cpml",
                cpml_layers=4,
                materials={"deb": (eps_inf, 0.01, 1.0), "d": (3.0, 0.0, 1.0)},
                debye={"deb": [(delta_eps, tau)]},
                geometry=geometry,
                ports=[((0.003, 0.006, 0.005), "ez", 0.0, o.Pulse(8e9, 0.8), True),
                       ((0.002, 0.003, 0.003), "ex", 50.0, o.Pulse(7e9, 0.9), True)],
                probes=[((0.013, 0.006, 0.005), "ez"), ((0.007, 0.006, 0.004), "ey"),
                        ((0.010, 0.005, 0.004), "hx")],
                lane=lane)


def lorentz_cavity_case(n_steps=6000):
    """PEC cavity 20 x 16 x 3 mm (dx 1 mm) filled with a lossless Lorentz
    medium (eps_inf 2, d_eps 1.5, f0 6 GHz, no damping), an Ez pulse off
    centre and an Ez probe; returns (spec, n_steps, the two TM110 polariton
    frequencies of the discrete dispersion relation)."""
    from tests.test_oracle_lorentz import polariton_freqs
    a, b, c, dx = 0.020, 0.016, 0.003, 1e-3
    w0 = 2 * np.pi * 6e9
    kappa = 1.5 * w0 ** 2
    spec = dict(freq_max=20e9, domain=(a, b, c), dx=dx, boundary="pec",
                materials={"lor": (2.0, 0.0, 1.0)}, lorentz={"lor": [(w0, 0.0, kappa)]},
                geometry=[((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), "lor")],
                ports=[((0.006, 0.005, 0.001), "ez", 0.0, o.Pulse(7e9, 1.0), True)],
                probes=[((0.013, 0.011, 0.001), "ez")], lane="run")
    g = o.UniformGrid(spec["freq_max"], spec["domain"], dx, 0)
    return spec, n_steps, polariton_freqs(2.0, w0, kappa, a, b, dx, g.dt)


def box_total(fluxes):
    """Net outward flux of one box from its six monitor spectra (lo face
    first, then hi face, per axis)."""
    return sum((1 if q % 2 else -1) * f for q, f in enumerate(fluxes))


def graded_xy_cavity_case(n_steps=8000):
    """Air PEC cavity graded in x (4:1) and y (2:1), uniform z: TM110 depends
    on both graded extents (tests/test_nonuniform_xy_cavity_accuracy.py of
    the reference)."""
    dx = 1e-3
    px = smooth_grading([dx] * 17 + [0.25e-3] * 8 + [dx] * 17, 1.3)
    py = smooth_grading([dx] * 15 + [0.5e-3] * 6 + [dx] * 15, 1.3)
    a, b, d = float(px.sum()), float(py.sum()), 30e-3
    f110 = (o.C0 / 2) * np.sqrt((1 / a) ** 2 + (1 / b) ** 2)
    spec = dict(freq_max=2 * f110, domain=(a, b, d), dx=dx, dx_profile=px, dy_profile=py,
                boundary="pec",
                ports=[((a / 3, b / 3, d / 2), "ez", 0.0, o.Pulse(f110, 0.8), True)],
                probes=[((2 * a / 3, 2 * b / 3, d / 2), "ez")])
    return spec, n_steps, f110


def mixed_graded_xy_case():
    """The mixed graded case with x and y graded too (and z graded)."""
    spec = mixed_graded_case()
    spec["dx_profile"] = smooth_grading([1e-3] * 5 + [0.5e-3] * 4 + [1e-3] * 7, 1.3)
    spec["dy_profile"] = smooth_grading([1e-3] * 4 + [0.4e-3] * 5 + [1e-3] * 5, 1.3)
    return spec


def patch_dz():
    """Graded dz of the patch workload: 10 x 1 mm, 13 cells 1 -> 0.1 mm
    (ratio 1.194), 16 x 0.1 mm substrate, 13 cells 0.1 -> 1 mm, 68 x 1 mm."""
    r = 0.1 ** (1.0 / 13.0)
    return np.asarray([1e-3] * 10 + [1e-3 * r ** k for k in range(1, 14)] + [1e-4] * 16
                      + [1e-4 / r ** k for k in range(1, 14)] + [1e-3] * 68)


def patch_case(n_xy=60):
    """The graded patch workload on a smaller x/y extent (oracle spec)."""
    dx = 5e-4
    dz = patch_dz()
    L = n_xy * dx
    z0 = float(np.sum(dz[:23]))
    z1 = z0 + 16 * 1e-4
    c = L / 2
    return dict(freq_max=4e9, domain=(L, L), dx=dx, dz_profile=dz, boundary="cpml",
                cpml_layers=12,
                materials={"fr4": (4.4, 0.002, 1.0), "pec": (1.0, 1e10, 1.0)},
                geometry=[((0.0, 0.0, z0), (L, L, z1), "fr4"),
                          ((0.0, 0.0, z0), (L, L, z0), "pec"),
                          ((c - 0.006, c - 0.007, z1), (c + 0.006, c + 0.007, z1), "pec")],
                ports=[((c - 0.003, c, z0), "ez", 50.0, o.Pulse(2.5e9, 1.2), True, 1.6e-3)],
                probes=[((c + 0.008, c, z1 + 2e-3), "ez")],
                s11_freqs=np.linspace(1e9, 4e9, 21))


def tile_case(lane="run"):
    """A uniform case larger than one tile of the temporal-blocked march in
    y and z (57 x 47 x 92 cells: four tile rows, two tile columns): lossy,
    magnetic and PEC boxes across tile edges, soft sources on three
    components, a lumped port, E and H probes on and off tile edges, DFT
    planes normal to x and z."""
    return dict(freq_max=12e9, domain=(0.040, 0.030, 0.075), dx=1e-3, boundary="cpml",
                cpml_layers=8,
                materials={"d": (3.5, 0.03, 1.0), "m": (1.0, 1e8, 1.0), "mag": (2.0, 0.01, 1.7),
                           "hi": (9.0, 0.0, 1.0)},
                geometry=[((0.006, 0.004, 0.010), (0.020, 0.016, 0.040), "d"),
                          ((0.022, 0.002, 0.050), (0.0235, 0.028, 0.0515), "m"),
                          ((0.010, 0.018, 0.045), (0.030, 0.026, 0.070), "mag"),
                          ((0.025, 0.008, 0.020), (0.035, 0.013, 0.030), "hi")],
                ports=[((0.012, 0.010, 0.030), "ez", 0.0, o.Pulse(9e9, 0.8), True),
                       ((0.030, 0.020, 0.055), "ex", 0.0, o.Pulse(7e9, 0.6), True),
                       ((0.005, 0.025, 0.065), "ey", 0.0, o.Pulse(10e9, 0.9), True),
                       ((0.028, 0.004, 0.015), "ez", 50.0, o.Pulse(8e9, 0.7), True)],
                probes=[((0.020, 0.015, 0.038), "ez"), ((0.004, 0.004, 0.004), "ez"),
                        ((0.015, 0.020, 0.052), "hx"), ((0.035, 0.027, 0.070), "ey"),
                        ((0.018, 0.003, 0.060), "hz")],
                dft_planes=[("z", 0.036, "ez", None, 4), ("x", 0.013, "hy", None, 3)],
                s11_freqs=np.array([7e9, 9e9]), lane=lane)


def tb_interior_case(lane="run"):
    """A uniform case with tiles of the temporal-blocked march that miss every
    CPML slab (43 x 34 x 130 cells, 4-layer CPML: the tile of rows 10-21 and
    columns 60-119 runs its middle x segments without CPML code) beside
    tiles that touch the slabs: boxes, sources, probes and a DFT plane on
    both kinds of tile."""
    return dict(freq_max=15e9, domain=(0.034, 0.025, 0.121), dx=1e-3, boundary="cpml",
                cpml_layers=4,
                materials={"d": (4.0, 0.02, 1.0), "mag": (1.5, 0.0, 2.5), "m": (1.0, 1e8, 1.0)},
                geometry=[((0.008, 0.007, 0.060), (0.026, 0.017, 0.090), "d"),
                          ((0.004, 0.002, 0.010), (0.012, 0.022, 0.030), "mag"),
                          ((0.020, 0.009, 0.100), (0.021, 0.016, 0.110), "m")],
                ports=[((0.017, 0.012, 0.075), "ez", 0.0, o.Pulse(10e9, 0.8), True),
                       ((0.010, 0.020, 0.020), "ey", 0.0, o.Pulse(8e9, 0.7), True)],
                probes=[((0.020, 0.012, 0.095), "ez"), ((0.017, 0.010, 0.070), "hy"),
                        ((0.030, 0.004, 0.015), "ex")],
                dft_planes=[("y", 0.012, "ez", None, 3)], lane=lane)
