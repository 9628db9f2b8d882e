"""The C-ABI library loads and exports every entry point include/fmm4rb.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import fmm4rb_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "fmm4rb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fmm4rb_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "fmm4rb_field" in names and "fmm4rb_field_device" in names
    assert set(names) == set(fmm4rb_hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(fmm4rb_hip.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_version_and_error_strings():
    assert "gfx950" in fmm4rb_hip.version()
    assert isinstance(fmm4rb_hip.lib().fmm4rb_last_error(), bytes)


def test_null_context_is_rejected():
    L = fmm4rb_hip.lib()
    assert L.fmm4rb_set_debug(None, 1) == 1
    assert b"NULL" in L.fmm4rb_last_error()
    assert L.fmm4rb_destroy(None) == 0
