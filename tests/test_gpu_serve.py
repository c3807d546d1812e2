"""Serving (continuous batching): every request's image is the same as when
its scene is rendered alone, while requests of different kinds and sizes
share batches and batches overlap on the device; admission respects the
HBM budget."""
import numpy as np
import pytest

from rusterizer_amd import _lib as L
from rusterizer_amd import scenes

pytestmark = pytest.mark.gpu


def _catalog():
    from rusterizer_amd.render import Context
    from rusterizer_amd.serve import Catalog

    root = Context(64, 64)
    kinds = {
        "batch": scenes.batch_scenes(n_scenes=12, seed=201, tri_lo=500, tri_hi=20000, resolutions=(128, 256, 384),
                                     tex_size=64),
        "soup": [scenes.soup_scene(40_000, 640, 360, seed=202)],
        "terrain": [scenes.terrain_scene(quads=150, width=480, height=270, tex_size=256)],
    }
    return root, Catalog(root, kinds)


def _alone(cat):
    """Each template rendered alone (a batch of one scene) on the root context."""
    import torch

    from rusterizer_amd import _lib as L

    out = {}
    for i, t in enumerate(cat.templates):
        img = torch.empty(t.pixels, dtype=torch.int32, device="cuda")
        row = np.array([t.row], dtype=L.batch_item_dtype())
        row[0]["argb_dev"] = img.data_ptr()
        cat.root.batch_begin()
        cat.root.batch_add(row)
        cat.root.batch_submit()
        cat.root.batch_wait()
        out[i] = img
    return out


@pytest.mark.parametrize("service", [False, True], ids=["own_raster", "raster_service"])
def test_serving_mix_equals_alone(service):
    """service: the slots post their rasters to the group's persistent raster
    service instead of launching them."""
    import torch

    from rusterizer_amd.serve import Server, make_requests

    root, cat = _catalog()
    try:
        alone = _alone(cat)
        reqs = make_requests(cat, 400, rate=40_000.0, weights={"batch": 0.8, "soup": 0.1, "terrain": 0.1}, seed=7)
        srv = Server(cat, slots=3, max_batch=12, service=service)
        try:
            rep = srv.run(reqs, keep_outputs=True)
            if service:
                assert root.service_stats()["launches"] >= 1
        finally:
            for r in reqs:
                if r.out is not None:
                    assert torch.equal(r.out, alone[r.tmpl]), (r.rid, cat.templates[r.tmpl].kind)
            srv.close()
        assert rep["requests"] == 400 and rep["rejected"] == 0 and rep["service"] == service
        assert set(rep["kinds"]) == {"batch", "soup", "terrain"}
        by_kind = rep["latency_ms_by_kind"]
        assert {k: v["n"] for k, v in by_kind.items()} == rep["kinds"]
        assert all(0.0 < v["p50"] <= v["p99"] for v in by_kind.values())
        assert rep["batches"] >= 400 // 12
        assert rep["max_batches_in_flight"] >= 2  # batches overlapped on the device
        mixed = {}
        for r in reqs:
            mixed.setdefault(r.batch, set()).add(cat.templates[r.tmpl].kind)
        assert any(len(k) >= 2 for k in mixed.values())  # kinds shared batches
    finally:
        cat.destroy()
        root.close()


@pytest.mark.parametrize("service", [False, True], ids=["own_raster", "raster_service"])
def test_serving_full_size_catalog(service):
    """Serving at its real size: the 256 batched scenes (seed 1002), the
    1M-triangle 1080p soup and the 4M-triangle 4K terrain resident, Poisson
    arrivals at 2000/s (seed 1004, 90% / 7% / 3%), 300 requests with their
    outputs kept, with the slots' rasters launched by the slots or run by
    the group's persistent raster service. Every output equals its template
    rendered alone, and the rendered-alone images of the terrain and of
    sampled batch scenes equal the oracle's (the 1M soup frame is checked
    against the oracle by test_soup_full_size)."""
    import torch

    from harness import run_oracle
    from rusterizer_amd.render import Context
    from rusterizer_amd.serve import Catalog, Server, make_requests

    kinds = {"batch": scenes.batch_scenes(), "soup": [scenes.soup_scene()], "terrain": [scenes.terrain_scene()]}
    root = Context(64, 64)
    cat = Catalog(root, kinds)
    try:
        reqs = make_requests(cat, 300, rate=2000.0, weights={"batch": 0.90, "soup": 0.07, "terrain": 0.03},
                             seed=scenes.SEEDS["serve"])
        srv = Server(cat, slots=3, max_batch=64, service=service)
        try:
            rep = srv.run(reqs, keep_outputs=True)
        finally:
            srv.close()
        assert rep["requests"] == 300 and rep["rejected"] == 0
        assert set(rep["kinds"]) == {"batch", "soup", "terrain"}, rep["kinds"]
        used = sorted({r.tmpl for r in reqs})
        alone = {}
        for i in used:  # each template used, rendered alone (a batch of one scene)
            t = cat.templates[i]
            img = torch.empty(t.pixels, dtype=torch.int32, device="cuda")
            row = np.array([t.row], dtype=L.batch_item_dtype())
            row[0]["argb_dev"] = img.data_ptr()
            cat.root.batch_begin()
            cat.root.batch_add(row)
            cat.root.batch_submit()
            cat.root.batch_wait()
            alone[i] = img
        for r in reqs:  # the first mismatch fails the test
            assert torch.equal(r.out, alone[r.tmpl]), (r.rid, cat.templates[r.tmpl].kind, r.batch)
        rng = np.random.default_rng(5)
        batch_used = [i for i in used if cat.templates[i].kind == "batch"]
        sample = [i for i in used if cat.templates[i].kind == "terrain"]
        sample += [batch_used[int(k)] for k in rng.choice(len(batch_used), 3, replace=False)]
        for i in sample:
            t = cat.templates[i]
            exp = run_oracle(t.scene)["img"]
            got = alone[i].cpu().numpy().view(np.uint32).reshape(t.scene.height, t.scene.width)
            print(f"serve full size: template {i} ({t.kind}, {t.scene.width}x{t.scene.height}) vs oracle", flush=True)
            assert np.array_equal(got, exp), (i, t.kind)
    finally:
        cat.destroy()
        root.close()


def test_serving_budget_admission():
    """Admission counts the device bytes the library reports (rz_memory,
    rz_batch_need): with a budget that holds the real need of the largest
    request but not the fixed-rate estimate of it, every request is served
    within the budget (batches are cut to fit and idle slots give memory
    back), and its images are right; a budget below the resident data
    rejects every request instead of waiting forever."""
    import torch

    from rusterizer_amd.serve import Server, make_requests

    root, cat = _catalog()
    try:
        alone = _alone(cat)
        probe = Server(cat, slots=3)
        try:
            fresh = probe.ctxs[0].memory()["held"]
            est = {i: probe.est_bytes(t) for i, t in enumerate(cat.templates)}
            dry = {i: probe.ctxs[0].batch_need(np.array([t.row], dtype=L.batch_item_dtype())) - fresh
                   for i, t in enumerate(cat.templates)}
        finally:
            probe.close()
        # what each template really took alone, against the library's up-front
        # estimate and the fixed-rate guess: they differ
        big = max(range(len(cat.templates)), key=lambda i: cat.templates[i].need)
        t_big = cat.templates[big]
        print("budget test: measured / library estimate / fixed guess of the largest:",
              t_big.need, dry[big], est[big], flush=True)
        assert any(abs(t.need - est[i]) > 0.2 * t.need for i, t in enumerate(cat.templates))
        rm = root.memory()
        base = rm["resident"] + rm["held"] + 3 * fresh
        headroom = cat.empty_bytes + t_big.need - fresh + 4 * t_big.pixels + (8 << 20)
        budget = base + headroom
        print(f"budget test: base {base}, headroom {headroom}, budget {budget}", flush=True)
        reqs = make_requests(cat, 150, rate=40_000.0, weights={"batch": 0.6, "soup": 0.2, "terrain": 0.2}, seed=8)
        srv = Server(cat, slots=3, budget_bytes=budget, max_batch=12)
        try:
            rep = srv.run(reqs, copy_back=True)
        finally:
            srv.close()
        assert rep["requests"] == 150 and rep["rejected"] == 0
        assert rep["peak_device_bytes"] <= budget
        assert rep["budget_cuts"] > 0  # some batches were cut down to fit
        for r in reqs:
            assert torch.equal(r.image, alone[r.tmpl].cpu()), r.rid
        # a request larger than the whole budget is rejected, not waited on forever
        reqs2 = make_requests(cat, 10, rate=40_000.0, weights={"terrain": 1.0}, seed=9)
        srv = Server(cat, slots=2, budget_bytes=base + 1000)
        try:
            rep2 = srv.run(reqs2)
        finally:
            srv.close()
        assert rep2["rejected"] == 10 and rep2["requests"] == 0
    finally:
        cat.destroy()
        root.close()


def test_cu_mask_same_results_and_errors():
    """A context restricted to a few CUs renders the same frame; the mask is
    refused while a frame is recorded and when it selects no CU."""
    import torch

    from harness import run_gpu
    from rusterizer_amd import scenes
    from rusterizer_amd.render import Context, RzError

    sc = scenes.soup_scene(20000, 300, 200, seed=91)  # colour FS: the context binds no texture
    ref = run_gpu(sc)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    ctx = Context(sc.width, sc.height)
    try:
        s0 = ctx.stream
        ctx.set_cu_mask(range(n_cu - 8, n_cu))
        assert ctx.stream != s0
        got = run_gpu(sc, ctx=ctx)
        assert np.array_equal(got["img"], ref["img"]) and np.array_equal(got["dep"], ref["dep"])
        with pytest.raises(RzError):
            ctx.set_cu_mask([n_cu + 40])  # no CU of the device
        m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
        ctx.draw(m, sc.fs)
        with pytest.raises(RzError):
            ctx.set_cu_mask(None)  # a frame is being recorded
        ctx.resolve()
        ctx.set_cu_mask(None)
        assert np.array_equal(run_gpu(sc, ctx=ctx)["img"], ref["img"])
        m.destroy()
    finally:
        ctx.close()


def test_serving_small_slot_on_reserved_cus():
    """With CUs reserved for small requests, every request still equals its
    image rendered alone, and no large request runs on the small slot."""
    import torch

    from rusterizer_amd.serve import Server, make_requests

    root, cat = _catalog()
    try:
        alone = _alone(cat)
        reqs = make_requests(cat, 300, rate=40_000.0, weights={"batch": 0.8, "soup": 0.1, "terrain": 0.1}, seed=11)
        small_tris = max(t.scene.ntri for t in cat.templates if t.kind == "batch")
        srv = Server(cat, slots=3, max_batch=12, small_cus=32, small_tris=small_tris)
        try:
            rep = srv.run(reqs, keep_outputs=True)
            for r in reqs:
                assert torch.equal(r.out, alone[r.tmpl]), (r.rid, cat.templates[r.tmpl].kind)
                if not srv.is_small(cat.templates[r.tmpl]):
                    assert r.slot != 0, r.rid
        finally:
            srv.close()
        assert rep["requests"] == 300 and rep["small_slot"]
        assert any(r.slot == 0 for r in reqs)
    finally:
        cat.destroy()
        root.close()
