"""ctypes binding of rfx_amd/librfxhip.so (the C ABI in include/rfxhip.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the soname libamdhip64.so.7; loading torch
first makes the dynamic linker reuse it instead of mapping a second HIP
runtime into the process).  There is no CPU fallback: if the library is
missing or no HIP device is visible, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must precede loading the HIP library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librfxhip.so")

RFXH_ZERO_EX = 1
RFXH_ZERO_EY = 2
RFXH_ZERO_EZ = 4
RFXH_POINT = 8
RFXH_ANISO = 16
RFXH_POLES = 0xFFFF0000
OP_PORT, OP_INJECT, OP_PROBE = 0, 1, 2

_P = C.c_void_p
_FP = C.POINTER(C.c_float)


class PointOp(C.Structure):
    _fields_ = [("cell", C.c_int64), ("kind", C.c_int32), ("comp", C.c_int32),
                ("slot", C.c_int32), ("order", C.c_int32), ("d_par", C.c_float),
                ("loop", C.c_float * 3)]


class Plane(C.Structure):
    _fields_ = [("axis", C.c_int32), ("index", C.c_int32), ("comp", C.c_int32),
                ("n_freq", C.c_int32), ("freqs", C.POINTER(C.c_double)),
                ("d_acc", _P)]


class Flux(C.Structure):
    _fields_ = [("axis", C.c_int32), ("index", C.c_int32), ("lo1", C.c_int32), ("n1", C.c_int32),
                ("lo2", C.c_int32), ("n2", C.c_int32), ("n_freq", C.c_int32),
                ("freqs", C.POINTER(C.c_double)), ("weight", _FP), ("d_acc", _P)]


class Ntff(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3), ("n_freq", C.c_int32),
                ("freqs", C.POINTER(C.c_double)), ("d_acc", _P * 6)]


class FarFieldDesc(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3), ("n_freq", C.c_int32),
                ("freqs", C.POINTER(C.c_double)), ("d_acc", _P * 6),
                ("n_axis", C.c_int32 * 3), ("pos", C.POINTER(C.c_double) * 3),
                ("width", C.POINTER(C.c_double) * 3), ("c0", C.c_double), ("eta0", C.c_double)]


class Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("graded", C.c_int32),
        ("inv_e", _FP * 3), ("inv_h", _FP * 3),
        ("dt", C.c_double),
        ("d_mat", _P), ("n_mat", C.c_int32),
        ("mat_coef", _FP), ("mat_flags", C.POINTER(C.c_uint32)),
        ("cpml_n", C.c_int32 * 6),
        ("cpml_b", _FP * 6), ("cpml_c", _FP * 6), ("cpml_km1", _FP * 6),
        ("cpml_inv_d", C.c_float * 6),
        ("pec_axes", C.c_int32),
        ("d_f", _P * 6), ("d_psi", _P), ("d_f2", _P * 6), ("d_psi2", _P),
        ("n_ops", C.c_int32), ("ops", C.POINTER(PointOp)),
        ("n_src", C.c_int32), ("d_src_wave", _P),
        ("n_probe", C.c_int32), ("d_probe_out", _P),
        ("n_port", C.c_int32), ("d_port_out", _P),
        ("max_steps", C.c_int64),
        ("n_plane", C.c_int32), ("planes", C.POINTER(Plane)),
        ("plane_t_offset", C.c_int32),
        ("n_flux", C.c_int32), ("flux", C.POINTER(Flux)),
        ("n_ntff", C.c_int32), ("ntff", C.POINTER(Ntff)),
        ("n_pole", C.c_int32), ("n_debye", C.c_int32), ("pole_coef", _FP), ("pol_lo", C.c_int32 * 3),
        ("pol_n", C.c_int32 * 3), ("d_pol", _P),
        ("aniso_coef", _FP),
    ]


class Region(C.Structure):
    _fields_ = [("i0", C.c_int32), ("j0", C.c_int32), ("k0", C.c_int32),
                ("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32)]


class AdjointDesc(C.Structure):
    _fields_ = [
        ("d_adj", _P * 6), ("d_adj_psi", _P), ("d_gbar", _P),
        ("d_probe_adj", _P), ("d_src_adj", _P),
        ("d_rec", _P), ("rec_step0", C.c_int64), ("rec_rows", C.c_int64),
        ("region", Region), ("d_grad", _P), ("dcoef", _FP),
        ("d_grad_sigma", _P), ("dcoef_sigma", _FP),
        ("d_rec_h", _P), ("d_grad_mu", _P), ("dcoef_mu", _FP),
        ("d_port_adj", _P), ("d_plane_adj", C.POINTER(_P)), ("d_flux_adj", C.POINTER(_P)),
        ("d_pol_adj", _P), ("d_rec_k", _P), ("dcoef_disp", _FP),
    ]


class AsmBox(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3), ("id", C.c_int32),
                ("pec", C.c_int32)]


_lib = None

EXPORTS = ("rfxh_psi_floats", "rfxh_plan_create", "rfxh_plan_destroy", "rfxh_run",
           "rfxh_state_set", "rfxh_set_state_set", "rfxh_set_timing", "rfxh_kernel_time_ms",
           "rfxh_run_record", "rfxh_run_record_h", "rfxh_run_record_ex", "rfxh_adjoint", "rfxh_batch_create",
           "rfxh_batch_destroy",
           "rfxh_batch_run", "rfxh_batch_run_record", "rfxh_batch_adjoint", "rfxh_set_variant",
           "rfxh_get_variant", "rfxh_ntff_far_field",
           "rfxh_ntff_far_field_noalloc", "rfxh_ntff_far_field_adjoint", "rfxh_assemble_fill",
           "rfxh_assemble_pec_pairs",
           "rfxh_assemble_pec_apply", "rfxh_last_error", "rfxh_version")


class RfxHipError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RfxHipError(
            f"{LIB_PATH} is missing: build it with `python -m rfx_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    lib.rfxh_psi_floats.argtypes = [C.POINTER(Desc)]
    lib.rfxh_psi_floats.restype = C.c_int64
    lib.rfxh_plan_create.argtypes = [C.POINTER(Desc), C.POINTER(_P)]
    lib.rfxh_plan_create.restype = C.c_int
    lib.rfxh_plan_destroy.argtypes = [_P]
    lib.rfxh_plan_destroy.restype = C.c_int
    lib.rfxh_run.argtypes = [_P, C.c_int64, C.c_int64, _P, C.c_int]
    lib.rfxh_run.restype = C.c_int
    lib.rfxh_state_set.argtypes = [_P]
    lib.rfxh_state_set.restype = C.c_int
    lib.rfxh_set_state_set.argtypes = [_P, C.c_int]
    lib.rfxh_set_state_set.restype = C.c_int
    lib.rfxh_set_timing.argtypes = [_P, C.c_int]
    lib.rfxh_set_timing.restype = C.c_int
    lib.rfxh_kernel_time_ms.argtypes = [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.rfxh_kernel_time_ms.restype = C.c_int
    lib.rfxh_run_record.argtypes = [_P, C.c_int64, C.c_int64, C.POINTER(Region), _P, _P]
    lib.rfxh_run_record.restype = C.c_int
    lib.rfxh_run_record_h.argtypes = [_P, C.c_int64, C.c_int64, C.POINTER(Region), _P, _P, _P]
    lib.rfxh_run_record_h.restype = C.c_int
    lib.rfxh_run_record_ex.argtypes = [_P, C.c_int64, C.c_int64, C.POINTER(Region), _P, _P, _P, _P]
    lib.rfxh_run_record_ex.restype = C.c_int
    lib.rfxh_adjoint.argtypes = [_P, C.POINTER(AdjointDesc), C.c_int64, C.c_int64, _P]
    lib.rfxh_adjoint.restype = C.c_int
    lib.rfxh_batch_create.argtypes = [C.POINTER(_P), C.c_int, C.POINTER(_P)]
    lib.rfxh_batch_create.restype = C.c_int
    lib.rfxh_batch_destroy.argtypes = [_P]
    lib.rfxh_batch_destroy.restype = C.c_int
    lib.rfxh_batch_run.argtypes = [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _P, C.c_int]
    lib.rfxh_batch_run.restype = C.c_int
    lib.rfxh_batch_run_record.argtypes = [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(Region), C.POINTER(_P), C.c_int, _P]
    lib.rfxh_batch_run_record.restype = C.c_int
    lib.rfxh_batch_adjoint.argtypes = [_P, C.POINTER(C.POINTER(AdjointDesc)),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64), _P]
    lib.rfxh_batch_adjoint.restype = C.c_int
    lib.rfxh_set_variant.argtypes = [_P, C.c_int, C.c_int]
    lib.rfxh_set_variant.restype = C.c_int
    lib.rfxh_get_variant.argtypes = [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.rfxh_get_variant.restype = C.c_int
    lib.rfxh_ntff_far_field.argtypes = [C.POINTER(FarFieldDesc), C.c_int32,
                                        C.POINTER(C.c_double), C.c_int32,
                                        C.POINTER(C.c_double), _P, _P, _P]
    lib.rfxh_ntff_far_field.restype = C.c_int
    lib.rfxh_ntff_far_field_noalloc.argtypes = lib.rfxh_ntff_far_field.argtypes
    lib.rfxh_ntff_far_field_noalloc.restype = C.c_int
    lib.rfxh_ntff_far_field_adjoint.argtypes = [C.POINTER(FarFieldDesc), C.c_int32,
                                                C.POINTER(C.c_double), C.c_int32,
                                                C.POINTER(C.c_double), _P, _P, C.POINTER(_P), _P]
    lib.rfxh_ntff_far_field_adjoint.restype = C.c_int
    lib.rfxh_assemble_fill.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(AsmBox),
                                       C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                                       C.c_int32, _P, _P, _P]
    lib.rfxh_assemble_fill.restype = C.c_int
    lib.rfxh_assemble_pec_pairs.argtypes = [C.c_int32, C.c_int32, C.c_int32, _P, _P, _P,
                                            C.c_int32, _P]
    lib.rfxh_assemble_pec_pairs.restype = C.c_int
    lib.rfxh_assemble_pec_apply.argtypes = [C.c_int32, C.c_int32, C.c_int32, _P, _P,
                                            C.POINTER(C.c_uint16), C.c_int32, _P]
    lib.rfxh_assemble_pec_apply.restype = C.c_int
    lib.rfxh_last_error.argtypes = []
    lib.rfxh_last_error.restype = C.c_char_p
    lib.rfxh_version.argtypes = []
    lib.rfxh_version.restype = C.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.rfxh_last_error().decode() if _lib is not None else "library not loaded"
        raise RfxHipError(f"rfxhip error {rc}: {msg}")


def require_device():
    if not torch.cuda.is_available():
        raise RfxHipError("no HIP device visible: the rfx_amd backend runs only on a GPU")
