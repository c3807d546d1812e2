"""ctypes binding of the CPU oracle (oracle/_build/librz_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py import this module to check or time the
reference algorithm. The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librz_oracle.so")

FS_TEXTURE, FS_COLOR, FS_DEBUG = 0, 1, 2
VS_MVP = 0
FLAG_TEXEL_OOB, FLAG_NO_TEXTURE, FLAG_CLIP_FEWER = 1, 2, 4

_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_u8 = C.POINTER(C.c_uint8)


class Setup(C.Structure):
    _fields_ = [
        ("pts", C.c_float * 6),
        ("normals", C.c_float * 6),
        ("depths", C.c_float * 3),
        ("depths_cs", C.c_float * 3),
        ("inv_2x_area", C.c_float),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with make (g++ -O2 -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "rzo_dot4": (C.c_float, [_f, _f]),
        "rzo_mat_mul": (None, [_f, _f, _f]),
        "rzo_mat_vec": (None, [_f, _f, _f]),
        "rzo_mvp": (None, [_f, _f, _f, _f]),
        "rzo_perspective_divide": (None, [_f, _f]),
        "rzo_viewport_setup": (None, [C.c_uint32, C.c_uint32, _f, C.POINTER(Setup)]),
        "rzo_setup_from_screen": (None, [_f, _f, C.POINTER(Setup)]),
        "rzo_eval_single": (None, [C.POINTER(Setup), C.c_float, C.c_float, _f]),
        "rzo_inside": (C.c_int, [C.POINTER(Setup), _f]),
        "rzo_coverage": (C.c_uint32, [C.POINTER(Setup), C.c_uint32, C.c_uint32, _f]),
        "rzo_fragment_depths": (None, [C.POINTER(Setup), C.c_uint32, _f, _f]),
        "rzo_interpolate": (None, [C.POINTER(Setup), _f, C.c_uint32, C.c_uint32, C.c_uint32, _f]),
        "rzo_mask_set": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_int]),
        "rzo_mask_get": (C.c_int, [C.c_uint32, C.c_uint32]),
        "rzo_try_clip": (C.c_int, [_f, _f, _f, _f, C.c_int, C.POINTER(C.c_int)]),
        "rzo_bbox": (None, [_f, _u64]),
        "rzo_box_filter": (C.c_uint32, [_u32]),
        "rzo_tiles_new": (C.c_void_p, [C.c_uint32, C.c_uint32]),
        "rzo_tiles_free": (None, [C.c_void_p]),
        "rzo_tiles_len": (C.c_uint32, [C.c_void_p]),
        "rzo_tiles_get": (None, [C.c_void_p, C.c_uint32, _u64]),
        "rzo_tiles_mark": (None, [C.c_void_p, C.c_uint64, C.c_uint64]),
        "rzo_tiles_next": (None, [C.c_void_p]),
        "rzo_tiles_marked": (C.c_uint32, [C.c_void_p]),
        "rzo_tiles_prev_marked": (C.c_uint32, [C.c_void_p]),
        "rzo_tiles_mask": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
        "rzo_to_argb": (C.c_uint32, [_f]),
        "rzo_from_rgba": (None, [_u8, _f]),
        "rzo_texture_sample": (None, [_u8, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _f, _u32]),
        "rzo_create": (C.c_void_p, [C.c_uint32, C.c_uint32]),
        "rzo_destroy": (None, [C.c_void_p]),
        "rzo_set_clear_depth": (None, [C.c_void_p, C.c_float]),
        "rzo_set_uniforms": (None, [C.c_void_p, _f, _f, _f]),
        "rzo_bind_texture": (C.c_int, [C.c_void_p, C.c_uint32, _u8, C.c_uint32, C.c_uint32]),
        "rzo_render": (C.c_int, [C.c_void_p, _f, _f, _f, C.c_uint32, _u32, C.c_uint64, C.c_int, C.c_int]),
        "rzo_render_mt": (C.c_int, [C.c_void_p, _f, _f, _f, C.c_uint32, _u32, C.c_uint64, C.c_int,
                                    C.c_int, C.c_int]),
        "rzo_read_msaa": (None, [C.c_void_p, _u32, _f]),
        "rzo_framebuffer": (None, [C.c_void_p, _u32]),
        "rzo_flags": (C.c_uint32, [C.c_void_p]),
        "rzo_stats": (None, [C.c_void_p, _u64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f)


def u32ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32)


def f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


# ---------------------------------------------------------------- pieces
def mat_mul(a, b) -> np.ndarray:
    a, b = f32(a).reshape(16), f32(b).reshape(16)
    out = np.zeros(16, np.float32)
    lib().rzo_mat_mul(fptr(a), fptr(b), fptr(out))
    return out.reshape(4, 4)


def mat_vec(m, v) -> np.ndarray:
    m, v = f32(m).reshape(16), f32(v).reshape(4)
    out = np.zeros(4, np.float32)
    lib().rzo_mat_vec(fptr(m), fptr(v), fptr(out))
    return out


def mvp(world, view, proj) -> np.ndarray:
    out = np.zeros(16, np.float32)
    lib().rzo_mvp(fptr(f32(world).reshape(16)), fptr(f32(view).reshape(16)),
                  fptr(f32(proj).reshape(16)), fptr(out))
    return out.reshape(4, 4)


def perspective_divide(clip) -> np.ndarray:
    c = f32(clip).reshape(12)
    out = np.zeros(12, np.float32)
    lib().rzo_perspective_divide(fptr(c), fptr(out))
    return out.reshape(3, 4)


def viewport_setup(w, h, ndc) -> Setup:
    s = Setup()
    lib().rzo_viewport_setup(w, h, fptr(f32(ndc).reshape(12)), C.byref(s))
    return s


def setup_from_screen(scr, w3) -> Setup:
    s = Setup()
    lib().rzo_setup_from_screen(fptr(f32(scr).reshape(9)), fptr(f32(w3).reshape(3)), C.byref(s))
    return s


def eval_single(s: Setup, x: float, y: float) -> np.ndarray:
    e = np.zeros(3, np.float32)
    lib().rzo_eval_single(C.byref(s), x, y, fptr(e))
    return e


def inside(s: Setup, e) -> bool:
    return bool(lib().rzo_inside(C.byref(s), fptr(f32(e).reshape(3))))


def coverage(s: Setup, x: int, y: int):
    ev = np.zeros(12, np.float32)
    m = lib().rzo_coverage(C.byref(s), x, y, fptr(ev))
    return int(m), ev


def fragment_depths(s: Setup, mask: int, evals) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().rzo_fragment_depths(C.byref(s), mask, fptr(f32(evals).reshape(12)), fptr(out))
    return out


def interpolate(s: Setup, attrs, x: int, y: int, cov: int) -> np.ndarray:
    out = np.zeros(6, np.float32)
    lib().rzo_interpolate(C.byref(s), fptr(f32(attrs).reshape(18)), x, y, cov, fptr(out))
    return out


def try_clip(verts, attrs=None, max_tris: int = 64):
    v = f32(verts).reshape(12)
    a = f32(np.zeros(18) if attrs is None else attrs).reshape(18)
    ov = np.zeros(max_tris * 12, np.float32)
    oa = np.zeros(max_tris * 18, np.float32)
    n = C.c_int(0)
    r = lib().rzo_try_clip(fptr(v), fptr(a), fptr(ov), fptr(oa), max_tris, C.byref(n))
    k = min(n.value, max_tris)
    return r, ov[: k * 12].reshape(k, 3, 4), oa[: k * 18].reshape(k, 3, 6)


def bbox(pts):
    out = np.zeros(4, np.uint64)
    lib().rzo_bbox(fptr(f32(pts).reshape(6)), out.ctypes.data_as(_u64))
    return tuple(int(x) for x in out)


def box_filter(c) -> int:
    a = np.ascontiguousarray(np.asarray(c, dtype=np.uint32))
    return int(lib().rzo_box_filter(u32ptr(a)))


def to_argb(rgba) -> int:
    return int(lib().rzo_to_argb(fptr(f32(rgba).reshape(4))))


def texture_sample(tex: np.ndarray, u: float, v: float):
    tex = np.ascontiguousarray(tex, dtype=np.uint8)
    th, tw = tex.shape[:2]
    out = np.zeros(4, np.float32)
    flags = C.c_uint32(0)
    lib().rzo_texture_sample(tex.ctypes.data_as(_u8), tw, th, u, v, fptr(out), C.byref(flags))
    return out, int(flags.value)


class Tiles:
    """rasterizer/buffers.rs BufferTiles through the oracle."""

    def __init__(self, w, h):
        self.p = lib().rzo_tiles_new(w, h)

    def __del__(self):
        if getattr(self, "p", None):
            lib().rzo_tiles_free(self.p)

    def __len__(self):
        return lib().rzo_tiles_len(self.p)

    def tile(self, i):
        b = np.zeros(4, np.uint64)
        lib().rzo_tiles_get(self.p, i, b.ctypes.data_as(_u64))
        return tuple(int(x) for x in b)

    def mark(self, row, col):
        lib().rzo_tiles_mark(self.p, row, col)

    def next(self):
        lib().rzo_tiles_next(self.p)

    def marked(self):
        return lib().rzo_tiles_marked(self.p)

    def prev_marked(self):
        return lib().rzo_tiles_prev_marked(self.p)

    def mask(self, which, idx):
        return bool(lib().rzo_tiles_mask(self.p, which, idx))


# ---------------------------------------------------------------- frames
class OracleRenderer:
    """Whole draw path of the reference (Renderer::render + framebuffer)."""

    def __init__(self, width: int, height: int, clear_depth: float | None = None):
        self.width, self.height = width, height
        self.p = lib().rzo_create(width, height)
        if not self.p:
            raise ValueError("bad size")
        if clear_depth is not None:
            lib().rzo_set_clear_depth(self.p, clear_depth)

    def __del__(self):
        if getattr(self, "p", None):
            lib().rzo_destroy(self.p)
            self.p = None

    def set_uniforms(self, world, view, proj):
        lib().rzo_set_uniforms(self.p, fptr(f32(world).reshape(16)), fptr(f32(view).reshape(16)),
                               fptr(f32(proj).reshape(16)))

    def bind_texture(self, index: int, rgba: np.ndarray):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        th, tw = rgba.shape[:2]
        rc = lib().rzo_bind_texture(self.p, index, rgba.ctypes.data_as(_u8), tw, th)
        if rc:
            raise ValueError(f"bind_texture rc={rc}")

    def render(self, pos, color, uv, idx, fs: int, vs: int = VS_MVP, threads: int = 1):
        pos = f32(pos).reshape(-1)
        nv = pos.size // 3
        color = f32(color if color is not None else np.zeros((nv, 4))).reshape(-1)
        uv = f32(uv if uv is not None else np.zeros((nv, 2))).reshape(-1)
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint32).reshape(-1))
        if threads <= 1:
            rc = lib().rzo_render(self.p, fptr(pos), fptr(color), fptr(uv), nv, u32ptr(idx),
                                  idx.size, vs, fs)
        else:
            rc = lib().rzo_render_mt(self.p, fptr(pos), fptr(color), fptr(uv), nv, u32ptr(idx),
                                     idx.size, vs, fs, threads)
        if rc:
            raise ValueError(f"render rc={rc}")

    def read_msaa(self):
        n = self.width * self.height * 4
        col = np.zeros(n, np.uint32)
        dep = np.zeros(n, np.float32)
        lib().rzo_read_msaa(self.p, u32ptr(col), fptr(dep))
        return (col.reshape(self.height, self.width, 4), dep.reshape(self.height, self.width, 4))

    def framebuffer(self) -> np.ndarray:
        out = np.zeros(self.width * self.height, np.uint32)
        lib().rzo_framebuffer(self.p, u32ptr(out))
        return out.reshape(self.height, self.width)

    @property
    def flags(self) -> int:
        return int(lib().rzo_flags(self.p))

    def stats(self):
        out = np.zeros(5, np.uint64)
        lib().rzo_stats(self.p, out.ctypes.data_as(_u64))
        return dict(zip(["tris", "outside", "inside", "clipped", "subtris"], (int(x) for x in out)))
