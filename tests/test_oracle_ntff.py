"""Pins of the oracle's near-to-far-field transform (oracle/fdtd_np.py
_ntff, ntff_axis_tables, far_field):

* the face accumulators equal DFT-plane accumulators of the same fields
  (themselves pinned by a direct DFT of the probe series) times the phase
  of the box's time base: E at n dt, H at (n + 1/2) dt;
* one face cell carrying one component radiates the textbook field of a
  single current element, with J = n x H and M = -n x E derived by hand
  per face (a sign or a swapped component in the cross products fails);
* a short dipole has the sin(theta) pattern and 1.76 dBi directivity (the
  reference's own bounds, tests/test_farfield.py:108-180 there);
* the far-field power (1/eta) integral |E|^2 dOmega equals the net outward
  Poynting flux through closed flux-monitor boxes around the source, which
  fixes the absolute factors k / (4 pi) and eta0.
"""
import numpy as np

from oracle import fdtd_np as o
from tests import cases

ETA0 = float(np.sqrt(o.MU_0 / o.EPS_0))


def _small_case():
    spec = cases.flux_case()
    spec.pop("flux_monitors")
    spec["ntff"] = ((0.010, 0.011, 0.009), (0.020, 0.019, 0.021),
                    np.array((9e9, 11e9), np.float32), 2)
    return spec


def test_faces_equal_time_shifted_dft_planes():
    """Uniform lane: a DFT plane sums at t = (n+1) dt, so the box's E face
    is the plane times e^{+j w dt} and its H face the plane times
    e^{+j w dt/2}.  Checked on the x_lo face (Ey, Ez, Hy, Hz) and the z_hi
    face (Ex, Ey, Hx, Hy)."""
    spec = _small_case()
    g0, cfg0 = o.build_uniform(spec, 1)
    (lo, hi, fr), = cfg0.ntff
    x_lo_mm = spec["ntff"][0][0]
    z_hi_mm = spec["ntff"][1][2]
    spec["dft_planes"] = ([("x", x_lo_mm, c, fr, len(fr)) for c in ("ey", "ez", "hy", "hz")]
                          + [("z", z_hi_mm, c, fr, len(fr)) for c in ("ex", "ey", "hx", "hy")])
    n = 120
    g, cfg = o.build_uniform(spec, n)
    out = o.run_loop(cfg, n)
    w = 2 * np.pi * np.asarray(fr, np.float64) * cfg.dt
    shift = [np.exp(1j * w), np.exp(1j * w), np.exp(0.5j * w), np.exp(0.5j * w)]
    assert cfg.dft_planes[0][1] == lo[0] and cfg.dft_planes[4][1] == hi[2]
    for t in range(4):
        xp = out.dft_planes[t][:, lo[1]:hi[1], lo[2]:hi[2]] * shift[t][:, None, None]
        np.testing.assert_allclose(out.ntff[0]["x_lo"][..., t], xp, rtol=1e-12,
                                   atol=1e-12 * np.abs(xp).max())
        zp = out.dft_planes[4 + t][:, lo[0]:hi[0], lo[1]:hi[1]] * shift[t][:, None, None]
        np.testing.assert_allclose(out.ntff[0]["z_hi"][..., t], zp, rtol=1e-12,
                                   atol=1e-12 * np.abs(zp).max())
    assert np.abs(out.ntff[0]["x_lo"]).max() > 0


def _one_cell(face, comp, lo=(2, 3, 4), hi=(6, 8, 9), value=1.0 + 0.5j):
    faces = {name: np.zeros((1,) + shp + (4,), np.complex128)
             for name, shp in zip(o.NTFF_FACES, o.ntff_face_shapes(lo, hi))}
    faces[face][0, 1, 2, comp] = value
    return faces


def test_single_face_cell_radiates_a_current_element():
    """Hand-derived surface currents (outward normal n):
      x_hi, Hz = v:  J = n x H = x x z v = -y v
      y_lo, Ex = v:  M = -n x E = -(-y) x x v = -z v
      z_hi, Hx = v:  J = z x x v = +y v
      x_lo, Ez = v:  M = -(-x) x z v = -y v
    and a current element J (M) at p radiates
      E_theta = -j k/(4 pi) (M.phi + eta0 J.theta) dS e^{j k r.p},
      E_phi   =  j k/(4 pi) (M.theta - eta0 J.phi) dS e^{j k r.p}."""
    lo, hi = (2, 3, 4), (6, 8, 9)
    dx = 1e-3
    shape = (12, 12, 12)
    pos, width = o.ntff_axis_tables(shape, dx, cpml_lo=(1, 1, 1))
    f = 40e9
    k = 2 * np.pi * f / o.C0
    theta = np.array([0.3, 1.1, 2.0])
    phi = np.array([0.2, 1.7, 4.0])
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    th_hat = np.stack([np.cos(TH) * np.cos(PH), np.cos(TH) * np.sin(PH), -np.sin(TH)], -1)
    ph_hat = np.stack([-np.sin(PH), np.cos(PH), np.zeros_like(TH)], -1)
    r_hat = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    v = 1.0 + 0.5j
    # (face, component slot, J, M, cell position)
    cases_ = [
        ("x_hi", 3, (0, -1, 0), (0, 0, 0), (hi[0], lo[1] + 1, lo[2] + 2)),
        ("y_lo", 0, (0, 0, 0), (0, 0, -1), (lo[0] + 1, lo[1], lo[2] + 2)),
        ("z_hi", 2, (0, 1, 0), (0, 0, 0), (lo[0] + 1, lo[1] + 2, hi[2])),
        ("x_lo", 1, (0, 0, 0), (0, -1, 0), (lo[0], lo[1] + 1, lo[2] + 2)),
    ]
    for face, comp, J, M, cell in cases_:
        et, ep = o.far_field(_one_cell(face, comp, lo, hi, v), lo, hi, [f], pos, width,
                             theta, phi)
        p = np.array([pos[a][cell[a]] for a in range(3)])
        J = np.array(J, float) * v * dx * dx
        M = np.array(M, float) * v * dx * dx
        e = np.exp(1j * k * (r_hat @ p))
        exp_t = -1j * k / (4 * np.pi) * ((ph_hat @ M) + ETA0 * (th_hat @ J)) * e
        exp_p = 1j * k / (4 * np.pi) * ((th_hat @ M) - ETA0 * (ph_hat @ J)) * e
        np.testing.assert_allclose(et[0], exp_t, rtol=1e-12, atol=1e-12 * np.abs(exp_t).max())
        np.testing.assert_allclose(ep[0], exp_p, rtol=1e-12, atol=1e-12 * np.abs(exp_p).max())


def _dipole():
    spec, n = cases.dipole_ntff_case()
    g, cfg = o.build_uniform(spec, n)
    out = o.run_loop(cfg, n)
    (lo, hi, fr), = cfg.ntff
    pos, width = o.ntff_axis_tables(g.shape, g.dx, cpml_lo=(g.pad,) * 3)
    return out.ntff[0], lo, hi, fr, pos, width


def _directivity_dbi(et, ep, theta, phi):
    power = np.abs(et[0]) ** 2 + np.abs(ep[0]) ** 2
    dth, dph = np.gradient(theta), np.gradient(phi)
    p_rad = np.sum(power * np.sin(theta)[:, None] * dth[:, None] * dph[None, :])
    return 10 * np.log10(4 * np.pi * power.max() / p_rad)


def test_short_dipole_pattern_and_directivity():
    """The reference's test records, for this set-up, 1.84 dBi with a
    full-circle phi sampling and 4.78 dBi with phi = (0, pi/2) (its
    tests/test_farfield.py:157-170 docstring, two decimals); theory for a
    Hertzian dipole is 1.76 dBi."""
    faces, lo, hi, fr, pos, width = _dipole()
    theta = np.linspace(0.01, np.pi - 0.01, 37)
    phi = np.linspace(0.0, 2 * np.pi, 24, endpoint=False)
    et, ep = o.far_field(faces, lo, hi, fr, pos, width, theta, phi)
    d_dbi = _directivity_dbi(et, ep, theta, phi)
    phi2 = np.array([0.0, np.pi / 2])
    d2 = _directivity_dbi(*o.far_field(faces, lo, hi, fr, pos, width, theta, phi2), theta, phi2)
    print("directivity dBi", d_dbi, d2)
    assert abs(d_dbi - 1.84) <= 0.006 and abs(d2 - 4.78) <= 0.006
    assert abs(d_dbi - 1.76) < 0.75
    # E-plane: |E_theta| follows sin(theta), E_phi small
    a = np.abs(et[0, :, 0])
    a = a / a.max()
    assert abs(theta[np.argmax(a)] - np.pi / 2) < np.pi / 6
    np.testing.assert_allclose(a, np.sin(theta) / np.sin(theta).max(), atol=0.05)
    assert np.abs(ep[0, :, 0]).max() < 0.3 * np.abs(et[0, :, 0]).max()


def test_far_field_power_equals_box_flux():
    """Radiated power two ways on the same run: (1/eta0) integral of
    |E_theta|^2 + |E_phi|^2 over the sphere (midpoint rule) against the net
    outward flux sum Re(E1 H2* - E2 H1*) dA of two closed monitor boxes.

    The box's time base stamps the E the state holds after step n (E at
    (n+1) dt, which the flux monitors use) with n dt, one step early
    against H; that costs 5-10 % of the power at 25-37 cells per
    wavelength.  With the E faces re-timed by e^{-j w dt} the two agree to
    1.5 %, which pins the absolute factors: a factor 2 in k/(4 pi) or a
    missing eta0 fails by far."""
    spec = cases.ntff_case()
    n = 400
    g, cfg = o.build_uniform(spec, n)
    out = o.run_loop(cfg, n)
    fl = [o.flux_spectrum(out.flux[q], cfg.flux_area[q]) for q in range(12)]
    small, big = cases.box_total(fl[:6]), cases.box_total(fl[6:])
    (lo, hi, fr), = cfg.ntff
    pos, width = o.ntff_axis_tables(g.shape, g.dx, cpml_lo=(g.pad,) * 3)
    nt, npf = 90, 72
    theta = (np.arange(nt) + 0.5) * np.pi / nt
    phi = (np.arange(npf) + 0.5) * 2 * np.pi / npf

    def power(faces):
        et, ep = o.far_field(faces, lo, hi, fr, pos, width, theta, phi)
        return np.sum((np.abs(et) ** 2 + np.abs(ep) ** 2) * np.sin(theta)[None, :, None],
                      axis=(1, 2)) * (np.pi / nt) * (2 * np.pi / npf) / ETA0

    ref = 0.5 * (small + big)
    p_ref = power(out.ntff[0])
    w = 2 * np.pi * np.asarray(fr, np.float64) * cfg.dt
    retimed = {k: v.copy() for k, v in out.ntff[0].items()}
    for v in retimed.values():
        v[..., :2] *= np.exp(-1j * w)[:, None, None, None]
    p_fix = power(retimed)
    print("box flux", small, big, "far field", p_ref / ref, "re-timed", p_fix / ref)
    assert np.all(np.abs(p_fix - ref) <= 0.015 * ref)
    assert np.all((p_ref < ref) & (p_ref > 0.85 * ref))


def test_graded_positions_reduce_to_uniform():
    """A graded z axis of equal cells gives the uniform positions and
    widths (the cumulative-edge rule of rfx/farfield.py:436-441 against the
    (k - cpml_lo) dx rule), and a 2:1 step moves the later nodes by the
    extra width."""
    shape, dx, p = (12, 10, 9), 1e-3, 3
    u_pos, u_w = o.ntff_axis_tables(shape, dx, cpml_lo=(p, p, p))
    g_pos, g_w = o.ntff_axis_tables(shape, dx, dz=np.full(shape[2], dx, np.float32),
                                    cpml_lo=(p, p, p))
    for a in range(3):
        # the z profile is float32 (as the reference keeps dz): 5e-8 relative
        np.testing.assert_allclose(g_pos[a], u_pos[a], rtol=0, atol=1e-9)
        np.testing.assert_allclose(g_w[a], u_w[a], rtol=1e-7)
    dz = np.full(shape[2], dx)
    dz[5] = 2 * dx
    g_pos, _ = o.ntff_axis_tables(shape, dx, dz=dz, cpml_lo=(p, p, p))
    np.testing.assert_allclose(g_pos[2][:6], u_pos[2][:6], atol=1e-12)
    np.testing.assert_allclose(g_pos[2][6:] - u_pos[2][6:], dx, atol=1e-12)
