"""The raster service (rz_service_*): a persistent raster kernel that stays
resident and renders the batches that the front ends of several contexts
post to it through device memory. Every batch it renders must equal the
context's own raster bit for bit (images, MSAA colours and depths, flags),
whatever else is in flight, across relaunches after idle exits and across
growth replays."""
import os
import time

import numpy as np
import pytest

from harness import assert_parity, run_batch, run_oracle
from rusterizer_amd import scenes
from test_gpu_batch import _mixed_scenes

pytestmark = pytest.mark.gpu


def _group(n, wg_per_cu=2, idle_us=50_000):
    from rusterizer_amd.render import Context

    root = Context(64, 64)
    ctxs = [Context(64, 64, share=root) for _ in range(n)]
    root.service_start(wg_per_cu, idle_us)
    for c in ctxs:
        c.service_attach(True)
    return root, ctxs


def _close(root, ctxs):
    for c in ctxs:
        c.close()
    root.service_stop()
    root.close()


def _same(a, b, where):
    assert np.array_equal(a["img"], b["img"]), where
    assert a["flags"] == b["flags"], where
    if "col" in a:
        assert np.array_equal(a["col"], b["col"]), where
        assert np.array_equal(a["dep"].view(np.uint32), b["dep"].view(np.uint32)), where


def test_service_matches_own_raster():
    scs = _mixed_scenes()
    ref, _ = run_batch(scs)
    root, ctxs = _group(1)
    try:
        for rep in range(3):  # graph replays of the posting sequence
            res, stats = run_batch(scs, ctx=ctxs[0])
            assert stats["tris"] == sum(s.ntri for s in scs)
            for sc, a, b in zip(scs, ref, res):
                _same(a, b, f"rep {rep} {sc.name} {sc.width}x{sc.height}")
        for sc, r in list(zip(scs, res))[:4]:
            assert_parity(r, run_oracle(sc), where=f"service {sc.name}")
        st = root.service_stats()
        assert st["on"] and st["launches"] >= 1 and st["grid"] == 2 * _n_cu()
    finally:
        _close(root, ctxs)


def _n_cu():
    import torch

    return torch.cuda.get_device_properties(0).multi_processor_count


def test_service_contexts_in_flight_together():
    """Three attached contexts with batches of different sizes in flight at
    once, in three rounds with the submission order rotated: every scene
    equals its batch rendered alone by the context's own raster."""
    works = [scenes.batch_scenes(n_scenes=4, seed=141, tri_lo=20000, tri_hi=120000, resolutions=(768, 1024),
                                 tex_size=64),
             scenes.batch_scenes(n_scenes=9, seed=142, tri_lo=500, tri_hi=8000, resolutions=(256, 512), tex_size=64),
             [scenes.soup_scene(3000, 200, 120, seed=143)]]
    alone = [run_batch(w, msaa=False)[0] for w in works]
    root, ctxs = _group(3)
    try:
        for rnd in range(3):
            order = [(rnd + k) % 3 for k in range(3)]
            subs = {}
            for k in order:
                _, outs, meshes = run_batch(works[k], ctx=ctxs[k], msaa=False, submit_only=True)
                subs[k] = (outs, meshes)
            pending = set(order)
            t0 = time.monotonic()
            while pending:
                for k in list(pending):
                    if ctxs[k].batch_query():
                        pending.discard(k)
                assert time.monotonic() - t0 < 60, "service batches did not finish"
            for k in order:
                outs, meshes = subs[k]
                for i, (sc, o, r) in enumerate(zip(works[k], outs, alone[k])):
                    img = o["img"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width)
                    assert np.array_equal(img, r["img"]), (rnd, k, i)
                    assert ctxs[k].batch_scene_flags(i) == r["flags"], (rnd, k, i)
                for m in meshes:
                    m.destroy()
    finally:
        _close(root, ctxs)


def test_service_relaunch_after_idle_exit():
    """A service that exits after 0.3 ms without work is relaunched by the
    next batch, which still comes out right."""
    scs = scenes.batch_scenes(n_scenes=5, seed=151, tri_lo=1000, tri_hi=20000, resolutions=(256, 512), tex_size=64)
    ref, _ = run_batch(scs, msaa=False)
    root, ctxs = _group(1, wg_per_cu=1, idle_us=300)
    try:
        for rep in range(3):
            res, _ = run_batch(scs, ctx=ctxs[0], msaa=False)
            for sc, a, b in zip(scs, ref, res):
                _same(a, b, f"rep {rep} {sc.name}")
            t0 = time.monotonic()
            while root.service_stats()["running"]:
                time.sleep(0.005)
                assert time.monotonic() - t0 < 10, "the service did not exit when idle"
        assert root.service_stats()["launches"] >= 3
    finally:
        _close(root, ctxs)


def test_service_growth_replay():
    """A posted batch whose clipped triangles overflow the first estimate is
    not posted; the host grows the buffers, replays it, and the service
    renders the replay."""
    heavy = scenes.soup_scene(20000, 256, 256, seed=121, clip_frac=0.9)
    small = scenes.soup_scene(1000, 64, 64, seed=122)
    root, ctxs = _group(1)
    try:
        res, stats = run_batch([small, heavy], ctx=ctxs[0])
        assert stats["replays"] >= 1
        assert_parity(res[0], run_oracle(small), where="small")
        assert_parity(res[1], run_oracle(heavy), where="heavy")
    finally:
        _close(root, ctxs)


def test_service_errors_and_detach():
    from rusterizer_amd.render import Context, RzError

    root = Context(64, 64)
    c = Context(64, 64, share=root)
    more = []
    try:
        with pytest.raises(RzError):
            c.service_attach(True)  # not running
        for wg, idle in ((0, 1000), (5, 1000), (2, 0)):
            with pytest.raises(RzError):
                root.service_start(wg, idle)
        root.service_start(1, 20_000)
        with pytest.raises(RzError):
            root.service_start(1, 20_000)  # already running
        c.service_attach(True)
        more = [Context(64, 64, share=root) for _ in range(16)]
        for m in more[:15]:
            m.service_attach(True)
        with pytest.raises(RzError):
            more[15].service_attach(True)  # 16 entries per group
        os.environ["RZ_RASTER_STATS"] = "1"
        try:
            counting = Context(64, 64, share=root)
        finally:
            del os.environ["RZ_RASTER_STATS"]
        more.append(counting)
        more[14].service_attach(False)
        with pytest.raises(RzError):
            counting.service_attach(True)  # the counting raster runs itself

        scs = scenes.batch_scenes(n_scenes=3, seed=161, tri_lo=1000, tri_hi=5000, resolutions=(256,), tex_size=64)
        ref, _ = run_batch(scs, msaa=False)
        _, outs, meshes = run_batch(scs, ctx=c, msaa=False, submit_only=True)
        with pytest.raises(RzError):
            root.service_stop()  # a batch is in flight
        c.batch_wait()
        for sc, o, r in zip(scs, outs, ref):
            assert np.array_equal(o["img"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width), r["img"])
        for m in meshes:
            m.destroy()
        root.service_stop()
        assert not root.service_stats()["on"]
        with pytest.raises(RzError):
            run_batch(scs, ctx=c, msaa=False)  # attached, service stopped
        c.service_attach(False)
        res, _ = run_batch(scs, ctx=c, msaa=False)  # its own raster again
        for a, b in zip(ref, res):
            assert np.array_equal(a["img"], b["img"])
        # a context closed with a posted batch in flight waits for it
        root.service_start(2, 50_000)
        c.service_attach(True)
        _, outs, meshes = run_batch(scs, ctx=c, msaa=False, submit_only=True)
        c.close()
        torch_imgs = [o["img"].cpu().numpy() for o in outs]
        for sc, img, r in zip(scs, torch_imgs, ref):
            assert np.array_equal(img.view(np.uint32).reshape(sc.height, sc.width), r["img"])
    finally:
        for m in more:
            m.close()
        c.close()
        root.service_stop()
        root.close()
