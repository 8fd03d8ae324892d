"""Host rules of the reverse mode (CPU): the recompute-segment count keeps
record + checkpoints inside the memory budget, follows the sqrt(n) split
when memory is tight and refuses what cannot fit."""
import math

import pytest

from rfx_amd.adjoint import segment_count


def _need(T, k, per_step, state):
    return math.ceil(T / k) * per_step + (k * state if k > 1 else 0)


def test_one_segment_when_the_record_fits():
    assert segment_count(3000, 100, 10**6, 3000 * 100) == 1


def test_forced_segments():
    assert segment_count(3000, 100, 10**6, 1 << 40, forced=7) == 7
    assert segment_count(5, 100, 10**6, 1 << 40, forced=9) == 5


@pytest.mark.parametrize("T,per_step,state,budget", [
    (2000, 36 * 533**3, 4 * (6 * 533**3 + 24 * 10 * 533**2), 400 << 30),   # whole 533^3 grid
    (3000, 36 * 16384 * 8, 4 * 6 * 11_432_621, 5 << 30),                     # large region, tight
    (4000, 1000, 50_000, 1_000_000),
])
def test_segments_fit_the_budget(T, per_step, state, budget):
    k = segment_count(T, per_step, state, budget)
    assert k >= 2
    assert _need(T, k, per_step, state) <= budget
    # the smallest fitting k: one fewer segment would not fit
    assert k == 2 or _need(T, k - 1, per_step, state) > budget


def test_whole_grid_533_is_refused_below_the_sqrt_minimum():
    T, per_step, state = 2000, 36 * 533**3, 4 * 6 * 533**3
    k_opt = round(math.sqrt(T * per_step / state))
    floor = _need(T, k_opt, per_step, state)
    with pytest.raises(MemoryError):
        segment_count(T, per_step, state, int(0.9 * floor))
    k = segment_count(T, per_step, state, int(1.01 * floor))
    assert k <= k_opt + 1 and _need(T, k, per_step, state) <= 1.01 * floor


def test_sigma_derivative_table_matches_finite_differences():
    """d ca / d sigma and d cb / d sigma against float64 central differences
    of ca = (1-L)/(1+L), cb = (dt/eps)/(1+L), L = sigma dt / (2 eps)."""
    import numpy as np
    from rfx_amd.adjoint import sigma_derivative_table
    from rfx_amd.constants import EPS_0
    from rfx_amd.materials import MaterialMap
    m = MaterialMap((2, 2, 2))
    for e, sg in ((1.0, 0.0), (4.0, 0.02), (11.7, 0.5), (2.2, 3.0)):
        m.get_id(e, sg, 1.0)
    dt = 1.9e-12
    tab = sigma_derivative_table(m, dt).astype(np.float64)
    eps_r, sig, _, _ = m.props()

    def coef(e, s):
        eps = e * EPS_0
        L = s * dt / (2 * eps)
        return (1 - L) / (1 + L), (dt / eps) / (1 + L)

    for q in range(len(m.table)):
        e, s = float(eps_r[q]), float(sig[q])
        h = 1e-4 * max(s, 1.0)
        (a1, b1), (a0, b0) = coef(e, s + h), coef(e, s - h)
        np.testing.assert_allclose(tab[q], [(a1 - a0) / (2 * h), (b1 - b0) / (2 * h)],
                                   rtol=1e-5)


def test_mu_derivative_table_matches_finite_differences():
    """d ch / d mu_r against a float64 central difference of ch = dt/(mu_r mu0)."""
    import numpy as np
    from rfx_amd.adjoint import mu_derivative_table
    from rfx_amd.constants import MU_0
    from rfx_amd.materials import MaterialMap
    m = MaterialMap((2, 2, 2))
    for mu in (1.0, 2.0, 4.5, 100.0):
        m.get_id(2.0, 0.0, mu)
    dt = 1.9e-12
    tab = mu_derivative_table(m, dt).astype(np.float64)
    _, _, mu_r, _ = m.props()
    for q in range(len(m.table)):
        mu, h = float(mu_r[q]), 1e-5 * float(mu_r[q])
        fd = (dt / ((mu + h) * MU_0) - dt / ((mu - h) * MU_0)) / (2 * h)
        np.testing.assert_allclose(tab[q], fd, rtol=1e-5)


@pytest.mark.parametrize("debye,lorentz", [
    ([(3.0, 8e-12)], []),
    ([], [(2 * 3.14159e10, 1e9, 2.0)]),
    ([(1.5, 5e-12), (0.7, 2e-11)], [(2 * 3.14159e10, 1e9, 2.0)]),
])
def test_dispersive_derivative_tables_match_finite_differences(debye, lorentz):
    """The dispersive reverse mode's tables against float64 central
    differences of the forward's own coefficient function
    (rfx_amd.materials.dispersion_tables) in eps_r and sigma: d ca, d cb
    (and d ce for eps_r), and d cc / dx = K-factor * cc for every pole's cc
    (the Debye cc_d and the Lorentz cc_L), on pole-free and pole-carrying
    materials with and without loss."""
    import numpy as np
    from types import SimpleNamespace
    from rfx_amd.adjoint import dispersive_derivative_tables
    from rfx_amd.materials import POLE_SHIFT, MaterialMap, dispersion_tables
    n_pole = len(debye) + len(lorentz)
    all_poles = ((1 << n_pole) - 1) << POLE_SHIFT
    entries = [(1.0, 0.0, 0), (4.0, 0.02, all_poles), (2.5, 0.0, all_poles), (6.0, 0.3, 0),
               (3.0, 0.1, (1 << POLE_SHIFT))]
    dt = 1.9e-12

    def tables(idx, de=0.0, ds=0.0):
        m = MaterialMap((2, 2, 2))
        for q, (e, s, f) in enumerate(entries):
            m.get_id(e + (de if q == idx else 0.0), s + (ds if q == idx else 0.0), 1.0, f)
        coef, _, rows = dispersion_tables(m, dt, debye, lorentz)
        return m, coef.astype(np.float64), rows.astype(np.float64)

    m0, _, _ = tables(-1)
    prob = SimpleNamespace(mat=m0, poles=debye, lorentz=lorentz)
    d_eps, d_sig, d_k = (t.astype(np.float64) for t in dispersive_derivative_tables(prob, dt))
    nd = len(debye)
    for q in range(1, len(entries)):
        e, s, _ = entries[q]
        # steps well above the float32 rounding of the coefficients and well
        # below their scales of variation (eps_r itself; 2 eps_inf / dt ~ 30
        # S/m for sigma)
        for which, h in ((0, 2e-2 * e), (1, 0.5)):
            _, c1, r1 = tables(q, *((h, 0.0) if which == 0 else (0.0, h)))
            _, c0, r0 = tables(q, *((-h, 0.0) if which == 0 else (0.0, -h)))
            fd = (c1[q] - c0[q]) / (2 * h)
            tab = d_eps[q] if which == 0 else d_sig[q]
            nc = 3 if which == 0 else 2
            np.testing.assert_allclose(tab[:nc], fd[:nc], rtol=2e-3, atol=1e-3 * abs(fd[:nc]).max())
            # every cc: the Debye rows' column 2, the Lorentz rows' column 3
            _, cm, rm = tables(-1)
            for p in range(n_pole):
                col = 2 if p < nd else 3
                cc = rm[q, p, col]
                dcc = (r1[q, p, col] - r0[q, p, col]) / (2 * h)
                np.testing.assert_allclose(d_k[q, which] * cc, dcc, rtol=2e-3,
                                           atol=1e-6 * abs(cc) + 1e-30)
