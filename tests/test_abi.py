"""The C ABI library builds, loads, and exports every entry point that
include/rz.h declares (no GPU calls here)."""
import ctypes
import os
import re

from rusterizer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "rz.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rz_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared() == sorted(_lib.SIGNATURES)


def test_library_exports_all_symbols():
    L = ctypes.CDLL(_lib.lib_path())
    for name in declared():
        assert hasattr(L, name), name


def test_null_context_errors():
    L = _lib.load()
    assert L.rz_last_error(None) == b"null context"
    assert L.rz_destroy(None) == 0
    assert L.rz_draw(None, None, 0, 0) < 0
    assert L.rz_resolve(None, None) < 0
    assert L.rz_set_uniforms(None, None, None, None) < 0
    assert L.rz_frame_flags(None) == 0
    assert L.rz_set_cu_mask(None, None, 0) < 0
    h = ctypes.c_void_p()
    assert L.rz_create(0, 10, ctypes.byref(h)) < 0
    assert L.rz_create(20000, 10, ctypes.byref(h)) < 0


def test_create_rejects_unsupported_screens():
    """Screens that need more than 256 coarse bins of 256 px are refused
    before any device call."""
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.rz_create(16384, 16384, ctypes.byref(h)) == -5
    assert L.rz_create(8192, 8192, ctypes.byref(h)) == -5
    assert L.rz_create(4097, 4096, ctypes.byref(h)) == -5
    assert not h.value
