"""ctypes binding of librz.so (the HIP rasterizer's C ABI, include/rz.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "librz.so")

RZ_VS_MVP = 0
RZ_FS_TEXTURE, RZ_FS_COLOR, RZ_FS_DEBUG = 0, 1, 2
RZ_FLAG_TEXEL_OOB, RZ_FLAG_CLIP_FEWER = 1, 4
RZ_STAGE_GEOMETRY, RZ_STAGE_BINNING, RZ_STAGE_RASTER, RZ_NSTAGES = 0, 1, 2, 3

# rz_batch_item of include/rz.h as a numpy record (same layout as the C struct)
BATCH_ITEM = None


def batch_item_dtype():
    import numpy as np

    global BATCH_ITEM
    if BATCH_ITEM is None:
        BATCH_ITEM = np.dtype([("width", "<u4"), ("height", "<u4"), ("clear_depth", "<f4"), ("texture", "<u4"),
                               ("vertex_shader", "<i4"), ("fragment_shader", "<i4"), ("argb_dev", "<u8"),
                               ("mesh", "<u8"), ("world", "<f4", (16,)), ("view", "<f4", (16,)),
                               ("proj", "<f4", (16,))], align=True)
        assert BATCH_ITEM.itemsize == 232, BATCH_ITEM.itemsize
    return BATCH_ITEM


# every entry point of include/rz.h: name -> (restype, argtypes)
_vp = C.c_void_p
_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
SIGNATURES = {
    "rz_create": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "rz_create_shared": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "rz_destroy": (C.c_int, [_vp]),
    "rz_last_error": (C.c_char_p, [_vp]),
    "rz_stream": (_vp, [_vp]),
    "rz_set_cu_mask": (C.c_int, [_vp, _u32, C.c_uint32]),
    "rz_set_clear_depth": (C.c_int, [_vp, C.c_float]),
    "rz_set_uniforms": (C.c_int, [_vp, _f, _f, _f]),
    "rz_bind_texture": (C.c_int, [_vp, C.c_uint32, _u8, C.c_uint32, C.c_uint32]),
    "rz_texture_count": (C.c_uint32, [_vp]),
    "rz_mesh_create": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp, C.c_uint64, C.POINTER(_vp)]),
    "rz_mesh_create_device": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp, C.c_uint64, C.POINTER(_vp)]),
    "rz_mesh_destroy": (C.c_int, [_vp, _vp]),
    "rz_draw": (C.c_int, [_vp, _vp, C.c_int, C.c_int]),
    "rz_resolve": (C.c_int, [_vp, _vp]),
    "rz_resolve_device": (C.c_int, [_vp, _vp]),
    "rz_resolve_msaa": (C.c_int, [_vp, _vp, _vp, _vp]),
    "rz_frame_flags": (C.c_uint32, [_vp]),
    "rz_frame_stats": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "rz_batch_begin": (C.c_int, [_vp]),
    "rz_batch_scene": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, _vp, _u32]),
    "rz_batch_scene_msaa": (C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    "rz_batch_draw": (C.c_int, [_vp, C.c_uint32, _vp, _f, _f, _f, C.c_uint32, C.c_int, C.c_int]),
    "rz_batch_add": (C.c_int, [_vp, _vp, C.c_uint32, _u32]),
    "rz_batch_submit": (C.c_int, [_vp]),
    "rz_batch_query": (C.c_int, [_vp]),
    "rz_batch_wait": (C.c_int, [_vp]),
    "rz_batch_scene_flags": (C.c_int, [_vp, C.c_uint32, _u32]),
    "rz_batch_tuning": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int]),
    "rz_batch_get_tuning": (C.c_int, [_vp, _u32]),
    "rz_set_coarse_bin": (C.c_int, [_vp, C.c_uint32]),
    "rz_memory": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "rz_batch_need": (C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(C.c_uint64)]),
    "rz_trim": (C.c_int, [_vp]),
    "rz_set_fine_tile": (C.c_int, [_vp, C.c_uint32]),
    "rz_get_fine_tile": (C.c_uint32, [_vp]),
    "rz_service_start": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "rz_service_stop": (C.c_int, [_vp]),
    "rz_service_attach": (C.c_int, [_vp, C.c_int]),
    "rz_service_stats": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "rz_profile": (C.c_int, [_vp, C.c_int]),
    "rz_profile_read": (C.c_int, [_vp, _f, _u32]),
    "rz_render_noalloc": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, _f, _f, _f, _vp, _vp, _vp,
                                    C.c_uint32, _vp, C.c_uint64, _vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                    _vp, _u32, C.POINTER(C.c_char_p)]),
}

_lib = None


def lib_path() -> str:
    return LIB_PATH


def load():
    """Load librz.so; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    # PyTorch ships its own HIP runtime; in one process it must initialise
    # the GPU before this library's runtime does, or it finds no device.
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.init()
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
