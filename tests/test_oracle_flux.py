"""Pins of the oracle's flux monitors (oracle/fdtd_np.py _flux,
flux_spectrum, dft_window_weights): energy conservation through nested
closed boxes around a source in vacuum, and the window closed forms."""
import numpy as np

from oracle import fdtd_np as o
from tests import cases


def test_nested_boxes_carry_the_same_power():
    """Lossless vacuum, CPML far away: the net outward flux of a closed box
    around the source is the radiated power, whatever the box size
    (0.5 % here, set by the co-located flux's O(dx^2) error).  A flux
    monitor with one term of Re(E1 H2* - E2 H1*) sign-flipped fails."""
    spec = cases.flux_case()
    n = 400
    g, cfg = o.build_uniform(spec, n)
    out = o.run_loop(cfg, n)
    fl = [o.flux_spectrum(out.flux[q], cfg.flux_area[q]) for q in range(12)]
    small, big = cases.box_total(fl[:6]), cases.box_total(fl[6:])
    print(small, big)
    assert np.all(small > 0) and np.all(big > 0)
    assert np.all(np.abs(big - small) <= 5e-3 * small)

    def bad_spectrum(accs, dA):
        e1, e2, h1, h2 = accs
        return np.real(np.sum((e1 * np.conj(h2) + e2 * np.conj(h1)) * dA, axis=(-2, -1)))

    fb = [bad_spectrum(out.flux[q], cfg.flux_area[q]) for q in range(12)]
    bs, bb = cases.box_total(fb[:6]), cases.box_total(fb[6:])
    assert not (np.all(np.abs(bb - bs) <= 5e-3 * np.abs(bs)) and np.all(bs > 0))


def test_window_weights_closed_forms():
    """rect -> none; hann = sin^2(pi x); tukey flat in the middle, equal to
    hann's shape on its tapers, alpha >= 1 -> hann."""
    N = 101
    assert o.dft_window_weights(N, "rect") is None
    h = o.dft_window_weights(N, "hann")
    x = np.arange(N + 1) / (N - 1)
    np.testing.assert_allclose(h, np.sin(np.pi * x) ** 2, atol=1e-6)
    t = o.dft_window_weights(N, "tukey", 0.4)
    mid = (x >= 0.2) & (x <= 0.8)
    assert np.all(t[mid] == 1.0)
    lo = x < 0.2
    np.testing.assert_allclose(t[lo], np.sin(np.pi * x[lo] / 0.4) ** 2, atol=1e-6)
    np.testing.assert_allclose(o.dft_window_weights(N, "tukey", 1.0), h, atol=0)
