"""The allocation-free entry point (rz_render_noalloc) renders exactly what
rz_set_uniforms + rz_draw + rz_resolve_device render in a fresh context:
bit-identical frames on consecutive calls that share the scratch array (the
second smaller than the first), the frame flags, and its errors."""
import ctypes as C

import numpy as np
import pytest

from rusterizer_amd import _lib, scenes
from rusterizer_amd.render import Context
from tests.harness import channel_diff, run_oracle

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")


def _fp(a):
    return np.ascontiguousarray(a, np.float32).reshape(16).ctypes.data_as(C.POINTER(C.c_float))


def render_noalloc(sc, idx=None, nv=None):
    """One frame through rz_render_noalloc; returns (rc, image, flags, message)."""
    import torch

    L = _lib.load()
    dev = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    idx = sc.idx if idx is None else idx
    pos, col, uv = dev(sc.pos), dev(sc.color), dev(sc.uv)
    ix = dev(np.ascontiguousarray(idx, np.uint32).view(np.int32))
    tex = None
    if sc.texture is not None:
        tex = dev(np.ascontiguousarray(sc.texture).reshape(-1).view(np.int32))
    out = torch.full((sc.height * sc.width,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    flags = C.c_uint32(0xFFFF)
    msg = C.c_char_p()
    th, tw = (sc.texture.shape[0], sc.texture.shape[1]) if sc.texture is not None else (0, 0)
    keep = (sc.world, sc.view, sc.proj)
    rc = L.rz_render_noalloc(None, sc.width, sc.height, float(sc.clear_depth), _fp(keep[0]), _fp(keep[1]),
                             _fp(keep[2]), ptr(pos), ptr(col), ptr(uv),
                             sc.pos.shape[0] if nv is None else nv, ptr(ix), int(np.size(idx)), ptr(tex), tw, th,
                             _lib.RZ_VS_MVP, int(sc.fs), ptr(out), C.byref(flags), C.byref(msg))
    torch.cuda.synchronize()
    img = out.cpu().numpy().view(np.uint32).reshape(sc.height, sc.width)
    return rc, img, flags.value, (msg.value or b"").decode()


def render_existing(sc):
    """The same frame through a fresh context and rz_resolve_device."""
    import torch

    ctx = Context(sc.width, sc.height)
    try:
        ctx.set_clear_depth(sc.clear_depth)
        if sc.texture is not None:
            ctx.bind_texture(0, sc.texture)
        ctx.set_uniforms(sc.world, sc.view, sc.proj)
        m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
        ctx.draw(m, sc.fs)
        out = torch.full((sc.height * sc.width,), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        ctx.resolve_device(out.data_ptr())
        torch.cuda.synchronize()
        flags = ctx.frame_flags()
        m.destroy()
        return out.cpu().numpy().view(np.uint32).reshape(sc.height, sc.width), flags
    finally:
        ctx.close()


def check_same(sc, oracle_check=False):
    print(f"noalloc: {sc.name} {sc.width}x{sc.height} {sc.ntri} tris fs={sc.fs}", flush=True)
    rc, img, flags, msg = render_noalloc(sc)
    assert rc == 0, (rc, msg)
    ref, ref_flags = render_existing(sc)
    assert np.array_equal(img, ref), f"{sc.name}: {(img != ref).sum()} pixels differ"
    assert flags == ref_flags
    if oracle_check:
        c = run_oracle(sc)
        assert np.array_equal(img, c["img"]), channel_diff(img, c["img"])
    return img


def test_noalloc_equals_resolve_device_two_calls():
    """A larger textured frame, then a smaller one (fewer triangles, smaller
    screen) on the same scratch array, then a third of another shader."""
    # few near-plane crossings: those triangles fill the screen and each
    # puts an entry into every tile list
    big = scenes.soup_scene(30_000, 1280, 720, seed=21, textured=True, clip_frac=0.02)
    small = scenes.cube_scene(320, 240)
    img_big = check_same(big, oracle_check=True)
    img_small = check_same(small, oracle_check=True)
    assert (img_big != 0xFF191919).mean() > 0.2 and (img_small != 0xFF191919).mean() > 0.05
    dbg = scenes.soup_scene(3000, 640, 480, seed=5)
    dbg.fs = scenes.FS_DEBUG
    check_same(dbg)


def test_noalloc_limits_and_errors():
    """Bad indices and oversize frames fail with a clear error; the scratch
    array then still renders the next frame correctly."""
    sc = scenes.soup_scene(2000, 256, 256, seed=3)
    bad = sc.idx.copy()
    bad[17] = sc.pos.shape[0]  # one past the last vertex
    rc, _, _, msg = render_noalloc(sc, idx=bad)
    assert rc == -1 and "index" in msg
    # every triangle spans the whole 1024x1024 screen inside the NDC cube (no
    # clipping): its 256 coarse bins times 20000 triangles outgrow the
    # array's coarse list
    n = 20_000
    big = np.tile(np.array([[-1, -1, 0.5], [0, 1, 0.5], [1, -1, 0.5]], np.float32), (n, 1))
    z = np.repeat(np.linspace(0.9, -0.9, n, dtype=np.float32), 3)
    big[:, 2] = z
    from rusterizer_amd import hostmath as hm

    full = scenes.Scene("fullscreen", 1024, 1024, big, np.ones((3 * n, 4), np.float32), None,
                        np.arange(3 * n, dtype=np.uint32), hm.identity(), hm.identity(), hm.identity(),
                        scenes.FS_COLOR, None)
    rc, _, _, msg = render_noalloc(full)
    assert rc == -5 and "bin lists do not fit" in msg, (rc, msg)
    check_same(sc)
    # more triangles than the entry point takes: refused before any device work
    many = scenes.Scene("many", 64, 64, np.zeros((3, 3), np.float32), None, None,
                        np.zeros(3 * 262_145, np.uint32), hm.identity(), hm.identity(), hm.identity(),
                        scenes.FS_COLOR, None)
    rc, _, _, msg = render_noalloc(many)
    assert rc == -5 and "262144" in msg
