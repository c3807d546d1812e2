"""GPU parity: the HIP library against the CPU oracle on the same seeded
inputs. Depth and coverage must match bit for bit; resolved colour within
1 LSB per channel (the tests also report whether colour matched exactly)."""
import math

import numpy as np
import pytest

from rusterizer_amd import hostmath as hm
from rusterizer_amd import meshes, scenes
from tests.harness import assert_parity, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")


def check(sc, **kw):
    g = run_gpu(sc, **kw)
    c = run_oracle(sc)
    r = assert_parity(g, c, where=sc.name if not isinstance(sc, list) else sc[0].name)
    print(f"parity ok: exact colour={r['exact_colour']} stats={g['stats']}", flush=True)
    return g, c


def test_cube_textured():
    sc = scenes.cube_scene()
    g, c = check(sc)
    assert c["stats"]["clipped"] >= 2  # near plane cuts the cube
    assert (g["img"] != 0xFF191919).mean() > 0.05


def test_cube_textured_clear_depth_one():
    """The plain cube workload clears depth to 1.0 (the reference clears to
    f32::MAX): every visible sample lies inside the far plane, so the image
    equals the f32::MAX one; depths are checked against the oracle."""
    g1, _ = check(scenes.cube_scene(clear_depth=1.0))
    gm = run_gpu(scenes.cube_scene())
    assert np.array_equal(g1["img"], gm["img"])
    # the same samples are written; the others keep their clear depth
    fmax = np.finfo(np.float32).max
    untouched = gm["dep"] == fmax
    assert (g1["dep"][untouched] == 1.0).all()
    assert np.array_equal(g1["dep"][~untouched], gm["dep"][~untouched])


@pytest.mark.parametrize("fs", [scenes.FS_COLOR, scenes.FS_TEXTURE, scenes.FS_DEBUG])
def test_soup_small_all_shaders(fs):
    sc = scenes.soup_scene(3000, 640, 480, seed=7, textured=(fs == scenes.FS_TEXTURE))
    sc.fs = fs
    check(sc)


def test_soup_1080p_slice():
    sc = scenes.soup_scene(60_000, 1920, 1080, seed=scenes.SEEDS["soup"])
    check(sc)


def test_soup_full_size():
    """The 1M-triangle 1920x1080 soup, full frame."""
    sc = scenes.soup_scene(1_000_000, 1920, 1080, seed=scenes.SEEDS["soup"])
    check(sc)


def test_clear_depth_one():
    sc = scenes.soup_scene(4000, 512, 512, seed=11)
    sc.clear_depth = 1.0
    check(sc)


@pytest.mark.parametrize("fine_tile", [8, 16])  # 32 is the default every other frame test uses
def test_fine_tile_sides(fine_tile):
    """Lists of 8x8 and 32x32 screen tiles (the raster still works on 16x16
    tiles): same frames, on uneven screens and every coarse bin side, on the
    terrain and on a screen at the side limit."""
    sc = scenes.soup_scene(5000, 333, 217, seed=12, textured=True)
    for cb in (0, 32, 48, 64, 128, 256):
        check(sc, coarse_bin=cb, fine_tile=fine_tile)
    check(scenes.terrain_scene(quads=200, width=640, height=360, tex_size=512), fine_tile=fine_tile)
    wide = scenes.soup_scene(3000, 16384, 24, seed=13)
    check(wide, fine_tile=fine_tile)


def test_uneven_screen_and_bins():
    sc = scenes.soup_scene(5000, 333, 217, seed=12, textured=True)
    for cb in (32, 48, 64, 128, 256):
        check(sc, coarse_bin=cb)
    small = scenes.soup_scene(2000, 200, 150, seed=14)
    check(small, coarse_bin=16)


def test_too_many_bins_rejected():
    from rusterizer_amd.render import Context, RzError

    ctx = Context(333, 217)
    try:
        with pytest.raises(RzError):
            ctx.set_coarse_bin(16)  # 21 x 14 bins > 256
        with pytest.raises(RzError):
            ctx.set_coarse_bin(24)  # not a multiple of 16
        with pytest.raises(RzError):
            ctx.set_coarse_bin(272)  # more than 256 tiles of 16x16 per bin
    finally:
        ctx.close()


@pytest.mark.parametrize("w,h", [(16384, 24), (24, 16384)])
def test_widest_screens(w, h):
    """Screens at the 16384-pixel side limit (bbox corners use all 15 bits,
    64 coarse bins along one axis)."""
    check(scenes.soup_scene(20000, w, h, seed=15, textured=True))


@pytest.mark.parametrize("tw,th", [(1, 1), (3, 257), (257, 3), (1000, 7)])
def test_odd_texture_sizes(tw, th):
    """Textures of one texel and of extreme aspect ratios, with uv in [0, 1]
    and, for a second soup, in [-0.5, 1.5) (clamped low, row-wrapping or
    out-of-range reads high, as the reference's linear texel index does)."""
    rng = np.random.default_rng(tw * 1000 + th)
    tex = rng.integers(0, 256, (th, tw, 4), dtype=np.uint8)
    sc = scenes.soup_scene(3000, 160, 120, seed=16, textured=True, texture=tex)
    check(sc)
    wide = scenes.soup_scene(3000, 160, 120, seed=17, textured=True, texture=tex)
    wide.uv = (wide.uv * 2.0 - 0.5).astype(np.float32)
    check(wide)


@pytest.mark.parametrize("tw,th", [((1 << 24) + 1, 1), (1, (1 << 24) + 1)])
def test_texture_side_past_24_bits(tw, th):
    """A texture side of 2^24 + 1 texels takes the raster's exact texel
    path (64-bit index arithmetic) instead of the 32-bit fast path."""
    rng = np.random.default_rng(18)
    tex = rng.integers(0, 256, (th, tw, 4), dtype=np.uint8)
    sc = scenes.soup_scene(2000, 128, 96, seed=19, textured=True, texture=tex)
    check(sc)


def test_one_pixel_screen():
    sc = scenes.soup_scene(500, 1, 1, seed=13)
    check(sc)


def test_multi_draw_demo_frame():
    """The Demo scene of main.rs:93-104: cube and sphere drawn into one frame
    with their own world matrices, 1280x720, projection of main.rs:137-142."""
    W, H = 1280, 720
    view = hm.Camera().get_view_matrix()
    proj = hm.project(1.0, 200.0, F(H) / F(W), hm.FRAC_PI_2)
    tex = meshes.demo_checkerboard()
    out = []
    for t in (0.0, 0.7, 2.3):
        e = F(t)
        w0 = hm.rotate(e, e, F(0.0))
        w1 = hm.mat_mul(hm.rotate(e, F(0.0), hm.FRAC_PI_4), hm.translate(0.0, 3.0, 0.0))
        for fs in (scenes.FS_TEXTURE, scenes.FS_COLOR, scenes.FS_DEBUG):
            parts = []
            for m, w in ((meshes.cube(1.0), w0), (meshes.sphere(0.5), w1)):
                parts.append(scenes.Scene("demo", W, H, m.vertices, m.colors, m.uvs, m.indices, w, view, proj,
                                          fs, tex))
            out.append(check(parts))


def test_clip_test_scene():
    """ClipTest mode of main.rs:106-125: a triangle swept along the window border."""
    W, H = 1280, 720
    view = hm.Camera().get_view_matrix()
    proj = hm.project(1.0, 200.0, F(H) / F(W), hm.FRAC_PI_2)
    tri = meshes.triangle()
    for t in np.linspace(0.0, 6.0, 13):
        e = F(t)
        world = hm.chain(hm.rotate_z(e), hm.translate(F(7.3), 0.0, 0.0), hm.rotate_z(F(-e)))
        sc = scenes.Scene("cliptest", W, H, tri.vertices, tri.colors, tri.uvs, tri.indices, world, view, proj,
                          scenes.FS_COLOR)
        check(sc)


def test_big_triangles_force_growth():
    """Near-clipped, screen-filling triangles: many coarse-bin entries per
    triangle, so the first frame outgrows its estimates and is replayed."""
    sc = scenes.soup_scene(3000, 1024, 1024, seed=21, clip_frac=0.9)
    g, _ = check(sc)
    assert g["stats"]["bin_entries"] > 0


def test_degenerate_and_outside():
    W, H = 256, 256
    pos = np.array([
        [0, 0, 2], [0, 1, 2], [0, 0, 2],          # zero area: culled
        [50, 50, 2], [51, 50, 2], [50, 51, 2],    # far outside: rejected
        [-1, -1, 2], [1, -1, 2], [0, 1, 2],       # ordinary
        [-1, -1, 2], [0, 1, 2], [1, -1, 2],       # back-facing: no samples
    ], dtype=np.float32)
    col = np.tile(np.array([[0.2, 0.4, 0.6, 1.0]], np.float32), (pos.shape[0], 1))
    view = hm.Camera().get_view_matrix()
    proj = hm.project(1.0, 200.0, F(1.0), hm.FRAC_PI_2)
    sc = scenes.Scene("degenerate", W, H, pos, col, None, np.arange(12, dtype=np.uint32), hm.identity(),
                      view, proj, scenes.FS_COLOR)
    check(sc)


def test_empty_frame_and_empty_mesh():
    sc = scenes.soup_scene(0, 64, 48, seed=1)
    g, c = check(sc)
    assert (g["img"] == 0xFF191919).all()


def test_determinism_and_bin_invariance():
    sc = scenes.soup_scene(20_000, 800, 600, seed=5, textured=True)
    a = run_gpu(sc)
    b = run_gpu(sc)
    c = run_gpu(sc, coarse_bin=64)
    for k in ("img", "col", "dep"):
        assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], c[k])


def test_consecutive_frames_reset():
    """framebuffer() clears: a second frame equals a fresh context's frame."""
    from rusterizer_amd.render import Context

    s1 = scenes.soup_scene(2000, 300, 200, seed=31)
    s2 = scenes.soup_scene(2000, 300, 200, seed=32)
    ctx = Context(300, 200)
    try:
        run_gpu(s1, ctx=ctx)
        second = run_gpu(s2, ctx=ctx)
    finally:
        ctx.close()
    fresh = run_gpu(s2)
    for k in ("img", "col", "dep"):
        assert np.array_equal(second[k], fresh[k])


@pytest.mark.parametrize("dz", [0.0, -1e-6, 1e-6, -1e-4])
def test_zfight_redraw(dz):
    """A soup drawn twice into one frame, the second copy shifted in depth by
    dz (0: identical depths, which strict LESS must reject everywhere). The
    depth culls work closest to their margins here."""
    a = scenes.soup_scene(20_000, 640, 480, seed=41)
    b = scenes.soup_scene(20_000, 640, 480, seed=41)
    b.pos = (b.pos + np.array([0.0, 0.0, dz], np.float32)).astype(np.float32)
    b.color = np.ascontiguousarray(b.color[:, ::-1])
    check([a, b])


def test_wide_clipper_same_results(monkeypatch):
    """Clipping with 8 lanes per polygon (default) and with a full wave
    (RZ_CLIP_WIDE=1, the path for polygons that outgrow 8 lanes) give the
    same frame, equal to the oracle."""
    sc = scenes.soup_scene(30_000, 640, 360, seed=1013, clip_frac=0.6)
    narrow = run_gpu(sc)
    monkeypatch.setenv("RZ_CLIP_WIDE", "1")
    wide = run_gpu(sc)
    assert narrow["stats"]["clipped"] > 10_000
    assert np.array_equal(narrow["img"], wide["img"])
    assert np.array_equal(narrow["dep"].view(np.uint32), wide["dep"].view(np.uint32))
    assert_parity(wide, run_oracle(sc), where="wide clipper")


def test_nine_vertex_polygons(monkeypatch):
    """Polygons of 9 vertices (more than the 8 lanes the clipper gives a
    polygon at first, so each is clipped again by the whole wave) render the
    same frame as the all-wide clipper, equal to the oracle."""
    sc = scenes.nine_vertex_scene()
    narrow = run_gpu(sc)
    monkeypatch.setenv("RZ_CLIP_WIDE", "1")
    wide = run_gpu(sc)
    assert narrow["stats"]["clipped"] == sc.ntri
    assert np.array_equal(narrow["img"], wide["img"])
    assert np.array_equal(narrow["dep"].view(np.uint32), wide["dep"].view(np.uint32))
    assert_parity(narrow, run_oracle(sc), where="nine-vertex polygons")


def test_exact_paths_same_results(monkeypatch):
    """With RZ_EXACT_PATHS=1 setup proves nothing, so every prim takes the
    raster's exact paths (f32::clamp of each barycentric, bounds-checked
    texel reads with the out-of-range check per fragment): colour, textured
    and clipped soups render the same frames and flags, equal to the oracle."""
    cases = [scenes.soup_scene(20_000, 480, 270, seed=1021, clip_frac=0.3),
             scenes.soup_scene(20_000, 480, 270, seed=1022, clip_frac=0.3, textured=True)]
    fast = [run_gpu(sc) for sc in cases]
    monkeypatch.setenv("RZ_EXACT_PATHS", "1")
    for sc, f in zip(cases, fast):
        e = run_gpu(sc)
        assert np.array_equal(f["img"], e["img"]), sc.name
        assert np.array_equal(f["dep"].view(np.uint32), e["dep"].view(np.uint32)), sc.name
        assert f["flags"] == e["flags"], sc.name
        assert_parity(e, run_oracle(sc), where=f"exact paths {sc.name}")


def test_terrain_small():
    """The terrain generator at reduced size: near-plane slicing of a
    detailed height field, textured."""
    sc = scenes.terrain_scene(quads=160, width=640, height=360, tex_size=256)
    g, _ = check(sc)
    assert g["stats"]["clipped"] > 0.05 * sc.ntri
    assert (g["img"] != 0xFF191919).mean() > 0.2


def test_terrain_full_size():
    """The 4M-triangle 3840x2160 terrain with its 4096^2 texture, full frame."""
    sc = scenes.terrain_scene()
    check(sc)


def _quad(z, half, uv_max, tex):
    pos = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]], np.float32)
    uv = np.array([[0, 0], [uv_max, 0], [uv_max, uv_max], [0, uv_max]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)  # front-facing for this camera
    col = np.ones((4, 4), np.float32)
    return scenes.Scene(f"quad{z}", 160, 120, pos, col, uv, idx, hm.identity(), hm.Camera().get_view_matrix(),
                        hm.project(1.0, 200.0, F(120 / 160), hm.FRAC_PI_2), scenes.FS_TEXTURE, tex)


def test_oob_flag_counts_overwritten_fragments():
    """The reference shades every depth-passing fragment in order, so a
    texel read past the texture by a fragment that a later one overwrites
    still happens (and would panic there): the flag must be raised although
    no such texel reaches the image. A fragment that fails the depth test is
    never shaded and must not raise it."""
    tex = meshes.checkerboard(64, 8)
    far_oob = _quad(3.0, 1.0, 1.5, tex)   # uv up to 1.5: reads past the texture
    near_ok = _quad(2.0, 1.6, 0.5, tex)   # in front, covers it entirely
    far_alone = run_oracle(far_oob)
    assert (far_alone["img"] != 0xFF191919).sum() > 100 and far_alone["flags"] & 1
    g, c = check([far_oob, near_ok])
    assert c["flags"] & 1 and g["flags"] & 1
    g2, c2 = check([near_ok, far_oob])    # the far quad now fails LESS everywhere
    assert not (c2["flags"] & 1) and not (g2["flags"] & 1)
    assert np.array_equal(g["img"], g2["img"])


def _lattice_scene(W=64, flip=False, step=4.0, ox=0.625, oy=0.125):
    """Triangles tiling a region with every edge (vertical, horizontal and
    diagonal) running exactly through sample positions: identity matrices,
    so screen = (W (x+1)/2, W (1-y)/2) exactly."""
    n = int((W - 2) // step)
    pts, idx, cols = [], [], []
    rng = np.random.default_rng(5)
    for j in range(n):
        for i in range(n):
            x0, y0 = ox + step * i, oy + step * j
            quad = [(x0, y0), (x0 + step, y0), (x0 + step, y0 + step), (x0, y0 + step)]
            for tri in ((0, 1, 2), (0, 2, 3)):
                t = [quad[k] for k in tri]
                if flip:
                    t = [t[0], t[2], t[1]]
                base = len(pts)
                for sx, sy in t:
                    pts.append([sx / (W / 2) - 1.0, 1.0 - sy / (W / 2), 0.25 + 0.5 * rng.random()])
                c = rng.random(4).astype(np.float32)
                c[3] = 1.0
                cols += [c, c, c]
                idx += [base, base + 1, base + 2]
    pos = np.array(pts, np.float32)
    sc = scenes.Scene("lattice", W, W, pos, np.array(cols, np.float32), None, np.array(idx, np.uint32),
                      hm.identity(), hm.identity(), hm.identity(), scenes.FS_COLOR)
    return sc


def test_tie_rule_on_shared_edges():
    """Edges through sample points: every such sample goes to exactly one of
    the two triangles sharing the edge, by the tie rule (mod.rs:148-170)."""
    for flip in (False, True):
        sc = _lattice_scene(flip=flip)
        c = run_oracle(sc)
        if (c["img"] != 0xFF191919).mean() < 0.5:
            continue  # this winding faces away
        g, _ = check(sc)
        covered = (c["col"] != 0xFF191919)
        assert covered.mean() > 0.5
        return
    raise AssertionError("no winding of the lattice was front-facing")


def _nonfinite_scene(fs):
    """Triangles with NaN, +-inf and 1e30 positions, a vertex on the camera
    plane (w = 0), NaN colour and NaN / inf / out-of-range uv beside ordinary
    ones: whatever the reference's arithmetic makes of them (rejects, clips,
    empty boxes, NaN depths that never pass LESS), the GPU must make the same."""
    nan, inf = F(np.nan), F(np.inf)
    tris = [
        [[-0.5, -0.5, 2], [0.5, -0.5, 2], [0.0, 0.5, 2]],
        [[nan, -0.5, 2], [0.5, -0.4, 2], [0.1, 0.5, 2]],
        [[-0.6, -0.6, 2], [inf, -0.5, 2], [0.0, 0.6, 2]],
        [[-0.6, 0.6, 2], [-inf, -0.5, 2], [0.2, 0.4, 2]],
        [[-1e30, -0.4, 2], [1e30, -0.4, 2], [0.0, 1e30, 2]],
        [[-0.3, -0.3, 0], [0.3, -0.3, 2], [0.0, 0.3, 2]],
        [[-0.4, -0.2, 1.5], [0.4, -0.2, 1.5], [0.0, 0.4, 1.5]],
    ]
    tris = [[t[0], t[2], t[1]] for t in tris]  # front-facing for this camera
    pos = np.array(tris, F).reshape(-1, 3)
    pos[:, 2] = -pos[:, 2]
    n = pos.shape[0]
    col = np.tile(np.array([[0.9, 0.3, 0.1, 1.0]], F), (n, 1))
    col[4] = [nan, 0.5, 0.5, 1.0]
    uv = np.tile(np.array([[0.25, 0.75]], F), (n, 1))
    uv[7] = [nan, 0.5]
    uv[10] = [inf, 0.2]
    uv[19] = [1.5, -0.25]
    proj = hm.project(0.5, 50.0, F(1.0), hm.FRAC_PI_2)
    return scenes.Scene(f"nonfinite_fs{fs}", 96, 80, pos, col, uv, np.arange(n, dtype=np.uint32), hm.identity(),
                        hm.identity(), proj, fs, meshes.checkerboard(64, 8))


@pytest.mark.parametrize("fs", [scenes.FS_COLOR, scenes.FS_TEXTURE, scenes.FS_DEBUG])
def test_nonfinite_inputs(fs):
    g, c = check(_nonfinite_scene(fs))
    assert c["stats"]["clipped"] >= 3 and (c["img"] != 0xFF191919).sum() > 100
