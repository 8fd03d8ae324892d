"""CPU checks of the C ABI library: it loads and exports every symbol that
include/rfxhip.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rfxhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(rfxh_\w+)\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "rfxh_plan_create" in names and "rfxh_run" in names and len(names) >= 8


def test_library_exports_every_declared_symbol():
    from rfx_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librfxhip.so not built (run python -m rfx_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(_declared())


def test_version_and_error_string_without_device():
    from rfx_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librfxhip.so not built")
    lib = _lib.load()
    assert lib.rfxh_version() >= 1
    # a null descriptor is rejected before any device call
    assert lib.rfxh_plan_create(None, None) == -1
    assert b"null" in lib.rfxh_last_error()


def test_descriptor_layout_matches_header():
    """The ctypes mirror of rfxh_desc has the members the header lists."""
    from rfx_amd import _lib
    txt = open(os.path.join(ROOT, "include", "rfxhip.h")).read()
    body = txt[txt.index("typedef struct {\n  int32_t nx, ny, nz;"):txt.index("} rfxh_desc;")]
    names = re.findall(r"[\s\*]([a-z_0-9]+)(?:\[\d\])?;", body)
    names += ["nx", "ny"]
    ct = [f[0] for f in _lib.Desc._fields_]
    for n in ct:
        assert n in body, n


def test_far_field_entry_rejects_missing_arguments_without_device():
    from rfx_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librfxhip.so not built")
    lib = _lib.load()
    assert lib.rfxh_ntff_far_field(None, 0, None, 0, None, None, None, None) == -1
    assert b"far field" in lib.rfxh_last_error()


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Every ctypes mirror has the size and member offsets that gcc gives
    the header's structs (a member added, dropped or reordered on one side
    fails)."""
    import shutil
    import subprocess
    from rfx_amd import _lib
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    mirrors = {"rfxh_point_op": _lib.PointOp, "rfxh_plane": _lib.Plane, "rfxh_flux": _lib.Flux,
               "rfxh_ntff": _lib.Ntff, "rfxh_desc": _lib.Desc, "rfxh_region": _lib.Region,
               "rfxh_adjoint_desc": _lib.AdjointDesc, "rfxh_farfield_desc": _lib.FarFieldDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rfxhip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                             text=True).stdout.splitlines():
        a, b, v = ln.split()
        got[(a, b)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
