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
