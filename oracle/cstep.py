"""ctypes binding of the C oracle stepper (oracle/cstep.c) — TEST INFRASTRUCTURE ONLY.

``run_c(cfg, n_steps, double=True)`` takes the same ``StepConfig`` as
``oracle.fdtd_np.run_loop`` and returns the same ``RunOutput`` (time series,
port V/I samples and DFTs, DFT-plane accumulators, final fields).  The C
code restates the identical step; tests check C == numpy on small grids.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import fdtd_np as o

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle_cstep.so")

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_long)
_I = C.POINTER(C.c_int)


class _Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
        ("graded", C.c_int), ("dx", C.c_double),
        ("inv_e", _F * 3), ("inv_h", _F * 3),
        ("dt", C.c_double),
        ("eps_r", _F), ("sigma", _F), ("mu_r", _F),
        ("cpml_n", C.c_long), ("cpml_axes", C.c_int),
        ("prof_b", (_F * 2) * 3), ("prof_c", (_F * 2) * 3), ("prof_k", (_F * 2) * 3),
        ("face_d", (C.c_double * 2) * 3),
        ("pec_axes", C.c_int), ("pec_bits", C.POINTER(C.c_ubyte)),
        ("n_steps", C.c_long), ("record_order", C.c_int),
        ("n_src", C.c_int), ("src_cell", _L), ("src_comp", _I), ("src_wave", _F),
        ("n_probe", C.c_int), ("probe_cell", _L), ("probe_comp", _I),
        ("n_port", C.c_int), ("port_cell", _L), ("port_comp", _I),
        ("port_dpar", _D), ("port_loop", _D),
        ("n_plane", C.c_int), ("plane_axis", _I), ("plane_index", _L),
        ("plane_comp", _I), ("plane_nf", _I),
        ("plane_freqs", C.POINTER(_F)), ("plane_acc", C.POINTER(_D)),
        ("init_fields", C.POINTER(_D)),
        ("aniso_eps", _F * 3),
    ]


class _Out(C.Structure):
    _fields_ = [("ts", _D), ("port_v", _D), ("port_i", _D), ("fields", C.POINTER(_D))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = C.CDLL(_SO)
        _lib.oracle_run.argtypes = [C.POINTER(_Desc), C.POINTER(_Out), C.c_int]
        _lib.oracle_run.restype = C.c_int
    return _lib


def _p(a, typ):
    return a.ctypes.data_as(typ)


def _cell(shape, idx):
    return (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]


def run_c(cfg: o.StepConfig, n_steps: int, double: bool = True,
          init_fields=None, want_fields: bool = True):
    if cfg.debye_poles or cfg.lorentz_poles:
        raise NotImplementedError("the C stepper has no dispersion ADE: use fdtd_np.run_loop")
    lib = _load()
    keep = []

    def K(a):
        keep.append(a)
        return a

    d = _Desc()
    d.nx, d.ny, d.nz = cfg.shape
    sp = cfg.spacing
    if sp["kind"] == "uniform":
        d.graded = 0
        d.dx = sp["dx"]
    else:
        d.graded = 1
        for a in range(3):
            d.inv_e[a] = _p(K(np.ascontiguousarray(sp["inv_e"][a], np.float32)), _F)
            d.inv_h[a] = _p(K(np.ascontiguousarray(sp["inv_h"][a], np.float32)), _F)
    d.dt = cfg.dt
    d.eps_r = _p(K(np.ascontiguousarray(cfg.eps_r, np.float32)), _F)
    d.sigma = _p(K(np.ascontiguousarray(cfg.sigma, np.float32)), _F)
    d.mu_r = _p(K(np.ascontiguousarray(cfg.mu_r, np.float32)), _F)
    if cfg.aniso_eps is not None:
        for q in range(3):
            d.aniso_eps[q] = _p(K(np.ascontiguousarray(cfg.aniso_eps[q], np.float32)), _F)
    d.cpml_n = cfg.cpml_n
    d.cpml_axes = sum(1 << a for a in range(3) if "xyz"[a] in cfg.cpml_axes) if cfg.cpml_n else 0
    for (ax, side), (b, c, k) in cfg.cpml_profiles.items():
        d.prof_b[ax][side] = _p(K(np.ascontiguousarray(b, np.float32)), _F)
        d.prof_c[ax][side] = _p(K(np.ascontiguousarray(c, np.float32)), _F)
        d.prof_k[ax][side] = _p(K(np.ascontiguousarray(k, np.float32)), _F)
    for ax, (dl, dh) in cfg.cpml_face_d.items():
        d.face_d[ax][0] = dl
        d.face_d[ax][1] = dh
    d.pec_axes = sum(1 << a for a in range(3) if "xyz"[a] in cfg.pec_axes)
    if cfg.pec_mask is not None:
        mx, my, mz = o.tangential_edge_masks(cfg.pec_mask)
        bits = (mx.astype(np.uint8) | (my.astype(np.uint8) << 1) | (mz.astype(np.uint8) << 2))
        d.pec_bits = _p(K(np.ascontiguousarray(bits)), C.POINTER(C.c_ubyte))
    d.n_steps = n_steps
    d.record_order = 0 if cfg.record_order == "uniform" else 1
    ci = {c: i for i, c in enumerate(o.COMPS)}
    d.n_src = len(cfg.sources)
    if cfg.sources:
        d.src_cell = _p(K(np.array([_cell(cfg.shape, s[0]) for s in cfg.sources], np.int64)), _L)
        d.src_comp = _p(K(np.array([ci[s[1]] for s in cfg.sources], np.int32)), _I)
        w = np.stack([np.asarray(s[2][:n_steps], np.float32) for s in cfg.sources], axis=1)
        d.src_wave = _p(K(np.ascontiguousarray(w)), _F)
    d.n_probe = len(cfg.probes)
    if cfg.probes:
        d.probe_cell = _p(K(np.array([_cell(cfg.shape, p[0]) for p in cfg.probes], np.int64)), _L)
        d.probe_comp = _p(K(np.array([ci[p[1]] for p in cfg.probes], np.int32)), _I)
    d.n_port = len(cfg.ports_vi)
    if cfg.ports_vi:
        d.port_cell = _p(K(np.array([_cell(cfg.shape, p[0]) for p in cfg.ports_vi], np.int64)), _L)
        d.port_comp = _p(K(np.array([ci[p[1]] for p in cfg.ports_vi], np.int32)), _I)
        d.port_dpar = _p(K(np.array([p[3] for p in cfg.ports_vi], np.float64)), _D)
        d.port_loop = _p(K(np.array([list(p[4]) for p in cfg.ports_vi], np.float64).ravel()), _D)
    d.n_plane = len(cfg.dft_planes)
    accs = []
    if cfg.dft_planes:
        d.plane_axis = _p(K(np.array([p[0] for p in cfg.dft_planes], np.int32)), _I)
        d.plane_index = _p(K(np.array([p[1] for p in cfg.dft_planes], np.int64)), _L)
        d.plane_comp = _p(K(np.array([ci[p[2]] for p in cfg.dft_planes], np.int32)), _I)
        d.plane_nf = _p(K(np.array([len(p[3]) for p in cfg.dft_planes], np.int32)), _I)
        fr = (_F * len(cfg.dft_planes))()
        ac = (_D * len(cfg.dft_planes))()
        for q, (ax, _, _, freqs) in enumerate(cfg.dft_planes):
            fr[q] = _p(K(np.ascontiguousarray(freqs, np.float32)), _F)
            acc = np.zeros((len(freqs),) + o._plane_shape(cfg.shape, ax), np.complex128)
            accs.append(acc)
            ac[q] = acc.ctypes.data_as(_D)
        K(fr)
        K(ac)
        d.plane_freqs = fr
        d.plane_acc = ac
    if init_fields is not None:
        arrs = [K(np.ascontiguousarray(init_fields[c], np.float64)) for c in o.COMPS]
        ptrs = (_D * 6)(*[_p(a, _D) for a in arrs])
        K(ptrs)
        d.init_fields = ptrs
    out = _Out()
    ts = np.zeros((n_steps, max(len(cfg.probes), 1)), np.float64)
    pv = np.zeros((max(len(cfg.ports_vi), 1), n_steps), np.float64)
    pi = np.zeros_like(pv)
    out.ts = _p(ts, _D)
    out.port_v = _p(pv, _D)
    out.port_i = _p(pi, _D)
    fields = None
    if want_fields:
        fields = {c: np.zeros(cfg.shape, np.float64) for c in o.COMPS}
        fp = (_D * 6)(*[_p(fields[c], _D) for c in o.COMPS])
        K(fp)
        out.fields = fp
    rc = lib.oracle_run(C.byref(d), C.byref(out), 1 if double else 0)
    if rc != 0:
        raise MemoryError("C oracle allocation failed")
    ts = ts[:, :len(cfg.probes)]
    dt = cfg.dt
    tvec = np.arange(n_steps) * dt
    vd, idd, pvl, pil = [], [], [], []
    for p, port in enumerate(cfg.ports_vi):
        ph = np.exp(-1j * 2.0 * np.pi * np.outer(np.asarray(port[2], np.float64), tvec)) * dt
        vd.append(ph @ pv[p])
        idd.append(ph @ pi[p])
        pvl.append(pv[p].copy())
        pil.append(pi[p].copy())
    return o.RunOutput(ts, accs, pvl, pil, vd, idd, fields, {})
