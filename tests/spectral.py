"""Frequency estimation for ring-down signals (test helper).

Matrix pencil method (Y. Hua and T. K. Sarkar, "Matrix pencil method for
estimating parameters of exponentially damped/undamped sinusoids in noise",
IEEE Trans. ASSP 38(5), 1990): the poles z_i = exp((-a_i + j 2 pi f_i) dt)
of a sum of damped complex exponentials are the eigenvalues of the pencil
formed by two shifted sub-blocks of the signal's dominant right singular
vectors.  Used where the reference's tests call Harminv.
"""
from __future__ import annotations

import numpy as np


def matrix_pencil_freqs(y, dt, rel_tol=1e-9, max_order=80, pencil=None):
    y = np.asarray(y, np.float64)
    n = len(y)
    L = pencil or n // 3
    rows = n - L
    H = np.lib.stride_tricks.sliding_window_view(y, L + 1)[:rows]
    _, s, vh = np.linalg.svd(H, full_matrices=False)
    m = int(min(max_order, np.sum(s > rel_tol * s[0])))
    V = vh[:m].conj().T                     # (L+1, m)
    V1, V2 = V[:-1], V[1:]
    z = np.linalg.eigvals(np.linalg.pinv(V1) @ V2)
    f = np.angle(z) / (2.0 * np.pi * dt)
    decay = -np.log(np.abs(z)) / dt
    keep = f > 0
    order = np.argsort(f[keep])
    return f[keep][order], decay[keep][order]
