"""Batched rendering (rz_batch_*): many independent scenes of uneven size in
one job. Every scene must come out exactly as the oracle renders it alone
(depth, coverage and colour bit-exact), identical to the
plain one-scene path, and independent of the tuning, of graph replay and of
other jobs in flight on the device."""
import numpy as np
import pytest

from harness import CLEAR_COLOR, assert_parity, run_batch, run_gpu, run_oracle
from rusterizer_amd import scenes

pytestmark = pytest.mark.gpu


def _mixed_scenes():
    """Small scenes covering the shader, size and clipping cases."""
    out = scenes.batch_scenes(n_scenes=8, seed=71, tri_lo=50, tri_hi=6000, resolutions=(16, 64, 200, 256),
                              tex_size=64)
    out.append(scenes.soup_scene(3000, 333, 77, seed=72))                       # colour FS, ragged size
    dbg = scenes.soup_scene(2000, 97, 130, seed=73)
    dbg.fs = scenes.FS_DEBUG                                                    # debug FS
    out.append(dbg)
    out.append(scenes.cube_scene(800, 600))                                     # near plane cuts the cube
    one = scenes.soup_scene(200, 1, 1, seed=74)                                 # 1x1 target
    out.append(one)
    c1 = scenes.soup_scene(4000, 128, 96, seed=75)
    c1.clear_depth = 1.0                                                        # clear depth 1.0
    out.append(c1)
    return out


def test_batch_scenes_match_oracle():
    scs = _mixed_scenes()
    res, stats = run_batch(scs)
    assert stats["tris"] == sum(s.ntri for s in scs)
    for sc, r in zip(scs, res):
        cpu = run_oracle(sc)
        assert_parity(r, cpu, where=f"{sc.name} {sc.width}x{sc.height}")


def test_batch_full_size_workload():
    """The full batched workload (256 textured soups, 9,203,286 triangles,
    127,795,200 px, its own seed) rendered as one job at its real buffer
    sizes: every scene's resolved image equals the oracle's (tile-parallel on
    the host), and every 8th scene and the one with the most triangles also
    match it sample by sample (depth, coverage, MSAA colours)."""
    scs = scenes.batch_scenes()
    res, stats = run_batch(scs)
    assert len(scs) == 256 and stats["tris"] == sum(s.ntri for s in scs) == 9_203_286
    assert sum(s.width * s.height for s in scs) == 127_795_200
    pick = set(range(0, len(scs), 8)) | {max(range(len(scs)), key=lambda i: scs[i].ntri)}
    for i, sc in enumerate(scs):
        c = run_oracle(sc)
        where = f"batch scene {i} {sc.width}x{sc.height} {sc.ntri} tris"
        if i in pick:
            assert_parity(res[i], c, where=where)
        else:
            assert np.array_equal(res[i]["img"], c["img"]), where


def test_batch_many_tiny_scenes():
    """3000 scenes of 1x1 to 48x48 pixels and 1 to 60 triangles in one job
    (many scenes per coarse-binning chunk and per raster band); a sample of
    them against the oracle, and every image against its plain render."""
    scs = scenes.batch_scenes(n_scenes=3000, seed=93, tri_lo=1, tri_hi=60, resolutions=(1, 7, 16, 33, 48),
                              tex_size=16)
    res, stats = run_batch(scs, msaa=True)
    assert stats["tris"] == sum(s.ntri for s in scs)
    for i in range(0, len(scs), 97):
        assert_parity(res[i], run_oracle(scs[i], threads=1), where=f"tiny scene {i}")
    for i in range(1, len(scs), 211):
        plain = run_gpu(scs[i], msaa=False)
        assert np.array_equal(res[i]["img"], plain["img"]), i


def test_batch_multi_draw_scene_and_empty_scene():
    a = scenes.soup_scene(3000, 200, 150, seed=81)
    b = scenes.soup_scene(2500, 200, 150, seed=82, textured=True)
    empty = scenes.soup_scene(0, 40, 30, seed=83)
    c = scenes.soup_scene(1500, 64, 64, seed=84)
    res, _ = run_batch([a, empty, c], extra_draws={0: [b]})
    assert_parity(res[0], run_oracle([a, b]), where="two draws")
    assert (res[1]["img"] == CLEAR_COLOR).all()
    assert_parity(res[2], run_oracle(c), where="after empty scene")


def test_batch_equals_plain_path():
    scs = scenes.batch_scenes(n_scenes=16, seed=91, tri_lo=1000, tri_hi=40000, resolutions=(256, 512),
                              tex_size=128)
    res, _ = run_batch(scs, msaa=False)
    for sc, r in zip(scs, res):
        plain = run_gpu(sc, msaa=False)
        assert np.array_equal(r["img"], plain["img"]), sc.name
        assert r["flags"] == plain["flags"], sc.name


TUNINGS = [
    dict(bin_target=256, raster_per_cu=0, use_graph=False, fine_tile=16),
    dict(bin_target=16, raster_per_cu=0, use_graph=False, fine_tile=16),
    dict(bin_target=4, raster_per_cu=2, use_graph=False, fine_tile=16),
    dict(bin_target=64, raster_per_cu=1, use_graph=True, fine_tile=16),
    dict(bin_target=256, raster_per_cu=8, use_graph=True, fine_tile=16),
    dict(bin_target=64, raster_per_cu=0, use_graph=True, fine_tile=8),
    dict(bin_target=16, raster_per_cu=4, use_graph=False, fine_tile=8),
    dict(bin_target=256, raster_per_cu=0, use_graph=True, fine_tile=32),
    dict(bin_target=4, raster_per_cu=0, use_graph=False, fine_tile=32),
]


def test_batch_tuning_invariance():
    scs = scenes.batch_scenes(n_scenes=10, seed=101, tri_lo=500, tri_hi=30000, resolutions=(256, 768),
                              tex_size=64)
    ref = None
    for t in TUNINGS:
        res, _ = run_batch(scs, tuning=t, msaa=False)
        imgs = [r["img"] for r in res]
        if ref is None:
            ref = imgs
            continue
        for sc, a, b in zip(scs, ref, imgs):
            assert np.array_equal(a, b), (t, sc.name)


def test_graph_replays_match_eager():
    """The captured graph, replayed several times with eager frames and
    synchronisations in between, gives the eager result every time."""
    from rusterizer_amd.render import Context

    scs = scenes.batch_scenes(n_scenes=6, seed=111, tri_lo=2000, tri_hi=20000, resolutions=(256, 512),
                              tex_size=64)
    eager, _ = run_batch(scs, tuning=dict(bin_target=64, raster_per_cu=0, use_graph=False), msaa=False)
    ctx = Context(64, 64)
    try:
        other = scenes.soup_scene(5000, 300, 200, seed=112)
        other_ref = run_gpu(other, msaa=False)["img"]
        for rep in range(4):
            res, _ = run_batch(scs, ctx=ctx, tuning=dict(bin_target=64, raster_per_cu=4, use_graph=True),
                               msaa=False)
            for sc, a, b in zip(scs, eager, res):
                assert np.array_equal(a["img"], b["img"]), (rep, sc.name)
            # eager work of another context between replays
            assert np.array_equal(run_gpu(other, msaa=False)["img"], other_ref)
    finally:
        ctx.close()


def test_batch_growth_replay():
    """A batch whose clipped triangles overflow the first estimate is replayed
    with larger buffers and still matches the oracle."""
    heavy = scenes.soup_scene(20000, 256, 256, seed=121, clip_frac=0.9)
    small = scenes.soup_scene(1000, 64, 64, seed=122)
    res, stats = run_batch([small, heavy])
    assert stats["clipped"] > 20000 // 8 + 1024  # more than the first clip-slot estimate
    assert stats["replays"] >= 1
    assert_parity(res[0], run_oracle(small), where="small")
    assert_parity(res[1], run_oracle(heavy), where="heavy")


def test_shared_contexts_in_flight_together():
    """Two contexts over the same meshes and textures, each with a batch in
    flight: every scene equals its result alone."""
    from rusterizer_amd.render import Context, RzError

    x = scenes.batch_scenes(n_scenes=5, seed=131, tri_lo=5000, tri_hi=60000, resolutions=(512, 768), tex_size=64)
    y = scenes.batch_scenes(n_scenes=7, seed=132, tri_lo=1000, tri_hi=30000, resolutions=(256, 512), tex_size=64)
    alone_x, _ = run_batch(x, msaa=False)
    alone_y, _ = run_batch(y, msaa=False)
    a = Context(64, 64)
    b = Context(64, 64, share=a)
    try:
        ca, outs_a, meshes_a = run_batch(x, ctx=a, msaa=False, submit_only=True)
        cb, outs_b, meshes_b = run_batch(y, ctx=b, msaa=False, submit_only=True)
        with pytest.raises(RzError):
            meshes_a[0].destroy()  # still used by the batch in flight
        while not (a.batch_query() and b.batch_query()):
            pass
        for sc, o, r in zip(x, outs_a, alone_x):
            assert np.array_equal(o["img"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width), r["img"])
        for sc, o, r in zip(y, outs_b, alone_y):
            assert np.array_equal(o["img"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width), r["img"])
        for m in meshes_a + meshes_b:
            m.destroy()
    finally:
        b.close()
        a.close()


def test_batch_errors():
    from rusterizer_amd.render import Context, RzError

    ctx = Context(64, 64)
    try:
        sc = scenes.soup_scene(100, 32, 32, seed=141)
        m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
        import torch

        out = torch.empty(32 * 32, dtype=torch.int32, device="cuda")
        with pytest.raises(RzError):
            ctx.batch_submit()                      # nothing recorded
        ctx.batch_begin()
        with pytest.raises(RzError):
            ctx.batch_begin()                       # already recording
        with pytest.raises(RzError):
            ctx.draw(m, sc.fs)                      # plain draw inside a batch
        with pytest.raises(RzError):
            ctx.batch_scene(5000, 5000, 1.0, out.data_ptr())  # more than 256 coarse bins
        with pytest.raises(RzError):
            ctx.batch_draw(0, m, sc.world, sc.view, sc.proj, 0, sc.fs)  # no scene 0 yet
        k = ctx.batch_scene(32, 32, 1.0, out.data_ptr())
        with pytest.raises(RzError):
            ctx.batch_draw(k, m, sc.world, sc.view, sc.proj, 3, scenes.FS_TEXTURE)  # texture not bound
        with pytest.raises(RzError):
            ctx.batch_draw(k, m, sc.world, sc.view, sc.proj, 0, 7)  # unknown shader
        ctx.batch_draw(k, m, sc.world, sc.view, sc.proj, 0, sc.fs)
        ctx.batch_submit()
        ctx.batch_wait()
        assert ctx.batch_query()
        ref = run_gpu(sc, msaa=False)["img"]
        assert np.array_equal(out.cpu().numpy().view(np.uint32).reshape(32, 32), ref)
        with pytest.raises(RzError):
            ctx.batch_tuning(bin_target=0)
        m.destroy()
    finally:
        ctx.close()


def test_batch_add_array_records_like_single_calls():
    """rz_batch_add (one call for many scenes) records exactly what
    rz_batch_scene + rz_batch_draw do, and adds nothing when an item is bad."""
    import torch

    from rusterizer_amd import _lib as L
    from rusterizer_amd import batch
    from rusterizer_amd.render import Context, RzError

    scs = scenes.batch_scenes(n_scenes=6, seed=151, tri_lo=500, tri_hi=20000, resolutions=(256, 512), tex_size=64)
    ref, _ = run_batch(scs, msaa=False)
    ctx = Context(64, 64)
    try:
        items = batch.upload(ctx, scs)
        table = batch.item_table(items)
        batch.run(ctx, items, table)
        for sc, r, it in zip(scs, ref, items):
            assert np.array_equal(batch.image(it), r["img"]), sc.name
        bad = table.copy()
        bad[3]["texture"] = 10_000  # not bound
        ctx.batch_begin()
        with pytest.raises(RzError):
            ctx.batch_add(bad)
        first = ctx.batch_add(table[:2])  # nothing of the failed call was added
        assert first == 0
        ctx.batch_submit()
        ctx.batch_wait()
        assert ctx.frame_stats()["tris"] == scs[0].ntri + scs[1].ntri
        for r in items:
            r.mesh.destroy()
        assert table.dtype == L.batch_item_dtype()
    finally:
        ctx.close()


def test_memory_report_and_trim():
    """rz_memory / rz_batch_need report the bytes a context holds and will
    hold; rz_trim gives the pipeline buffers back and the next batch (which
    grows them again and re-captures its graph) renders the same images."""
    from rusterizer_amd import _lib as L
    from rusterizer_amd import batch
    from rusterizer_amd.render import Context

    scs = scenes.batch_scenes(n_scenes=5, seed=131, tri_lo=2000, tri_hi=30000, resolutions=(256, 512),
                              tex_size=64)
    ctx = Context(64, 64)
    try:
        items = batch.upload(ctx, scs)
        table = batch.item_table(items)
        m0 = ctx.memory()
        assert m0["resident"] >= sum(sc.pos.nbytes + sc.idx.nbytes + sc.uv.nbytes + sc.texture.nbytes for sc in scs)
        need = ctx.batch_need(table)
        assert need > m0["held"]
        batch.record(ctx, items, table)
        assert ctx.memory()["job"] == need  # the recorded batch, the same estimate
        ctx.batch_submit()
        ctx.batch_wait()
        first = [batch.image(r).copy() for r in items]
        m1 = ctx.memory()
        assert m1["held"] >= need  # the buffers the job sized (grown further only on overflow)
        assert ctx.batch_need(table) == m1["held"]  # the same batch needs no more
        ctx.trim()
        m2 = ctx.memory()
        assert m2["held"] < m1["held"] // 10 and m2["resident"] == m1["resident"]
        batch.run(ctx, items, table)
        for r, img in zip(items, first):
            assert np.array_equal(batch.image(r), img), r.scene.name
        for r in items:
            r.mesh.destroy()
        assert ctx.memory()["resident"] == m0["resident"] - sum(
            sc.pos.nbytes + sc.idx.nbytes + sc.uv.nbytes + (sc.color.nbytes if sc.color is not None else 0)
            for sc in scs)
    finally:
        ctx.close()
    _ = L
