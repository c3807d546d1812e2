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


def test_noalloc_rejects_bad_arguments_without_a_device():
    """rz_render_noalloc checks its arguments before any device call."""
    L = _lib.load()
    m = (ctypes.c_float * 16)(*([1.0, 0, 0, 0, 0] * 3 + [1.0]))
    msg = ctypes.c_char_p()
    flags = ctypes.c_uint32(7)
    dummy = 0x1000  # never dereferenced: each call fails its checks first

    def call(w=64, h=64, fs=1, vs=0, ni=3, nv=3, pos=dummy, idx=dummy, out=dummy, tex=None, tw=0, th=0):
        return L.rz_render_noalloc(None, w, h, 1.0, m, m, m, pos, None, None, nv, idx, ni, tex, tw, th, vs, fs, out,
                                   ctypes.byref(flags), ctypes.byref(msg))

    assert call(w=0) == -1 and b"width" in msg.value and flags.value == 0
    assert call(w=16384, h=16384) == -5
    assert call(out=None) == -1
    assert call(pos=None) == -1
    assert call(ni=4) == -1 and b"multiple of 3" in msg.value
    assert call(vs=1) == -1
    assert call(fs=3) == -1
    assert call(fs=0) == -3  # texture shader without a texture
    assert call(fs=0, tex=dummy, tw=65536, th=65536) == -5
    assert call(ni=3 * 262145) == -5 and b"262144" in msg.value
