"""The reference-shaped Python API (render.Renderer / Uniforms / Mesh /
Texture, mirroring src/render.rs:38-145 and src/uniform.rs:12-46) renders the
Demo frame of src/main.rs:93-104 exactly as the oracle does, keeps the
reference's texture-binding rule, and frees the device copies of meshes it
replaces."""
import gc

import numpy as np
import pytest

from rusterizer_amd import hostmath as hm
from rusterizer_amd import meshes, render, scenes
from tests.harness import channel_diff, run_oracle

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")


W, H = 1280, 720


def _demo_matrices(t):
    """main.rs:93-104, 137-142: camera, projection and the two world matrices."""
    e = F(t)
    view = hm.Camera().get_view_matrix()
    proj = hm.project(1.0, 200.0, F(H) / F(W), hm.FRAC_PI_2)
    w_cube = hm.rotate(e, e, F(0.0))
    w_sphere = hm.mat_mul(hm.rotate(e, F(0.0), hm.FRAC_PI_4), hm.translate(0.0, 3.0, 0.0))
    return view, proj, w_cube, w_sphere


def test_renderer_demo_frame_matches_oracle():
    """render.rs:98-114 through the mirror: write_block, then render per
    mesh, then framebuffer(); three shaders, two times of the animation."""
    r = render.Renderer(W, H)
    try:
        tex = meshes.demo_checkerboard()
        r.uniforms().bind_texture(0, render.Texture(tex))
        cm, sm = meshes.cube(1.0), meshes.sphere(0.5)
        cube = render.Mesh(cm.vertices, cm.indices, cm.colors, cm.uvs)
        sphere = render.Mesh(sm.vertices, sm.indices, sm.colors, sm.uvs)
        for t in (0.0, 0.7):
            view, proj, w0, w1 = _demo_matrices(t)
            for fs in (render.FragmentShader.TEXTURE, render.FragmentShader.COLOR, render.FragmentShader.DEBUG):
                blk = r.uniforms().write_block()
                blk.view, blk.projection = view, proj
                blk.world = w0
                r.render(cube, render.VertexShader.MVP, fs)
                blk.world = w1
                r.render(sphere, render.VertexShader.MVP, fs)
                img = r.framebuffer()
                parts = [scenes.Scene("demo", W, H, m.vertices, m.colors, m.uvs, m.indices, w, view, proj,
                                      int(fs), tex) for m, w in ((cm, w0), (sm, w1))]
                c = run_oracle(parts)
                print(f"renderer demo t={t} fs={fs!r}", flush=True)
                assert np.array_equal(img, c["img"]), (t, fs, channel_diff(img, c["img"]))
                assert (img != 0xFF191919).mean() > 0.02
        # framebuffer() resolves and clears: an empty frame is the clear colour
        assert (r.framebuffer() == 0xFF191919).all()
        assert r.device_meshes() == 2
    finally:
        r.close()


def test_renderer_texture_binding_rule():
    """uniform.rs:29-33 asserts index == textures.len()."""
    r = render.Renderer(64, 64)
    try:
        t = render.Texture(meshes.checkerboard(8, 2))
        with pytest.raises(ValueError):
            r.uniforms().bind_texture(1, t)
        r.uniforms().bind_texture(0, t)
        assert r.uniforms().get_texture(0) is t
        assert r.context.texture_count() == 1
    finally:
        r.close()


def test_renderer_replaced_meshes_are_freed():
    """A mesh edited and invalidated gets a new device copy; the old one is
    destroyed after the frame that may still use it, and so is the copy of
    a Mesh that was garbage-collected."""
    r = render.Renderer(256, 256)
    try:
        tri = meshes.triangle()
        m = render.Mesh(tri.vertices, tri.indices, tri.colors, tri.uvs)
        view, proj = hm.Camera().get_view_matrix(), hm.project(1.0, 200.0, F(1.0), hm.FRAC_PI_2)
        blk = r.uniforms().write_block()
        blk.view, blk.projection = view, proj
        r.render(m, render.VertexShader.MVP, render.FragmentShader.COLOR)
        a = r.framebuffer()
        m.vertices[:, 0] *= F(0.5)
        m.invalidate()
        r.render(m, render.VertexShader.MVP, render.FragmentShader.COLOR)
        assert r.device_meshes() == 2  # old copy retired, still owned by the pending frame
        b = r.framebuffer()
        assert r.device_meshes() == 1
        assert not np.array_equal(a, b)
        sc = scenes.Scene("tri", 256, 256, m.vertices, m.colors, m.uvs, m.indices, hm.identity(), view, proj,
                          scenes.FS_COLOR)
        assert np.array_equal(b, run_oracle(sc)["img"])
        for _ in range(5):  # meshes created per frame and dropped
            tmp = render.Mesh(tri.vertices, tri.indices, tri.colors, tri.uvs)
            r.render(tmp, render.VertexShader.MVP, render.FragmentShader.COLOR)
            del tmp
            gc.collect()
            r.framebuffer()
        assert r.device_meshes() == 1
    finally:
        r.close()
