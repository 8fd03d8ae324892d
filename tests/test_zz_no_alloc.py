"""The allocation-free far-field entry point (rfxh_ntff_far_field_noalloc)
against rfxh_ntff_far_field through the C ABI, on synthetic face
accumulators: two calls in a row (the second smaller, so stale tables of
the first would show), each to 1e-10 relative to the largest field value
(the no-alloc entry computes the direction vectors on the device: float64
rounding, not bit-identity); requests beyond its table limit are
refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rfx_amd import _lib
    return _lib


def _problem(rng, shape, lo, hi, nf):
    import torch
    from rfx_amd import _lib
    n = [hi[a] - lo[a] for a in range(3)]
    face_dims = [(n[1], n[2]), (n[1], n[2]), (n[0], n[2]), (n[0], n[2]), (n[0], n[1]),
                 (n[0], n[1])]
    accs = []
    for d1, d2 in face_dims:
        x = rng.standard_normal((4, nf, d1, d2, 2))
        accs.append(torch.from_numpy(x).to("cuda").view(torch.complex128).reshape(4, nf, d1, d2))
    pos = [np.ascontiguousarray(np.cumsum(rng.uniform(0.8e-3, 1.2e-3, shape[a]))) for a in range(3)]
    width = [np.ascontiguousarray(rng.uniform(0.8e-3, 1.2e-3, shape[a])) for a in range(3)]
    freqs = np.ascontiguousarray(np.linspace(8e9, 12e9, nf))
    d = _lib.FarFieldDesc()
    for a in range(3):
        d.lo[a], d.hi[a] = lo[a], hi[a]
        d.n_axis[a] = shape[a]
        d.pos[a] = pos[a].ctypes.data_as(C.POINTER(C.c_double))
        d.width[a] = width[a].ctypes.data_as(C.POINTER(C.c_double))
    d.n_freq = nf
    d.freqs = freqs.ctypes.data_as(C.POINTER(C.c_double))
    for q in range(6):
        d.d_acc[q] = accs[q].data_ptr()
    d.c0, d.eta0 = 299792458.0, 376.730313668
    keep = (accs, pos, width, freqs)
    return d, keep


def _call(lib_mod, fn, d, theta, phi, nf):
    import torch
    e_th = torch.zeros((nf, len(theta), len(phi)), dtype=torch.complex128, device="cuda")
    e_ph = torch.zeros_like(e_th)
    dp = C.POINTER(C.c_double)
    lib_mod.check(fn(C.byref(d), len(theta), theta.ctypes.data_as(dp), len(phi),
                     phi.ctypes.data_as(dp), e_th.data_ptr(), e_ph.data_ptr(),
                     torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return e_th.cpu().numpy(), e_ph.cpu().numpy()


def test_noalloc_far_field_matches_the_allocating_entry(lib):
    L = lib.load()
    rng = np.random.default_rng(7)
    cases = [((40, 36, 30), (5, 6, 4), (30, 28, 25), 5, 37, 24),
             ((22, 20, 18), (4, 3, 5), (14, 15, 12), 2, 9, 7)]
    for shape, lo, hi, nf, nt, npf in cases:
        _say(f"far field no-alloc, box {lo}-{hi} of {shape}, {nf} freqs, {nt}x{npf} directions")
        d, keep = _problem(rng, shape, lo, hi, nf)
        theta = np.ascontiguousarray(np.linspace(0.05, np.pi - 0.05, nt))
        phi = np.ascontiguousarray(np.linspace(0.0, 2 * np.pi, npf, endpoint=False))
        ref = _call(lib, L.rfxh_ntff_far_field, d, theta, phi, nf)
        got = _call(lib, L.rfxh_ntff_far_field_noalloc, d, theta, phi, nf)
        big = max(np.abs(ref[0]).max(), np.abs(ref[1]).max())
        assert big > 0
        for a, b in zip(got, ref):
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10 * big)


def test_noalloc_far_field_refuses_tables_beyond_its_limit(lib):
    L = lib.load()
    rng = np.random.default_rng(8)
    d, keep = _problem(rng, (12, 12, 12), (2, 2, 2), (9, 9, 9), 1)
    theta = np.ascontiguousarray(np.linspace(0.0, np.pi, 400))
    phi = np.ascontiguousarray(np.linspace(0.0, 2 * np.pi, 300))   # 9 x 120000 doubles > 2^20
    with pytest.raises(lib.RfxHipError, match="scratch array"):
        _call(lib, L.rfxh_ntff_far_field_noalloc, d, theta, phi, 1)
