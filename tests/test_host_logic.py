"""Host-side logic that needs no GPU: the bench's command-line contract and
roofline fields, the autotuner's candidate set, signature and cache, and the
serving loop's arrival process."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bench_roofline_fields_and_committed_traffic():
    import bench

    for name in ("traffic_raster_batch.json", "traffic_raster_soup1m.json", "traffic_raster_terrain.json"):
        t, src = bench.load_traffic(name)
        assert isinstance(t, int) and t > 0, name
        assert src.startswith("profiles/") and os.path.exists(os.path.join(ROOT, src)), src
    rf = bench.roofline(1_442_252_848, 8.0, bench.load_traffic("traffic_raster_batch.json"))
    assert set(rf) >= {"bound", "achieved", "peak", "unit", "frac", "traffic", "traffic_source"}
    assert rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["achieved"] - 1_442_252_848 / 8e-3 / 1e9) < 1e-6
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    assert bench.load_traffic("no_such_profile.json") == (None, None)


def test_autotune_candidates_signature_and_cache(tmp_path):
    from rusterizer_amd import autotune

    cands = list(autotune.candidates())
    assert len(cands) == len(autotune.FINE_TILES) * len(autotune.BIN_TARGETS) * len(autotune.RASTER_PER_CU)
    assert all(c["use_graph"] for c in cands)
    # the tile sides asked for (8x8 / 16x16 / 32x32) and the bin capacities are all searched
    assert {c["fine_tile"] for c in cands} == {8, 16, 32}
    assert {c["bin_target"] for c in cands} == set(autotune.BIN_TARGETS)
    # the signature depends on the resolution histogram and the triangle count's magnitude only
    a = autotune.signature([(256, 256), (512, 512), (256, 256)], 100_000)
    b = autotune.signature([(512, 512), (256, 256), (256, 256)], 120_000)
    c = autotune.signature([(512, 512), (512, 512), (256, 256)], 100_000)
    assert a == b and a != c
    path = str(tmp_path / "tune.json")
    cache = autotune.TuningCache(path)
    assert cache.get(a) is None
    cache.put(a, {"bin_target": 64, "raster_per_cu": 0, "use_graph": True})
    again = autotune.TuningCache(path)
    assert again.get(a) == {"bin_target": 64, "raster_per_cu": 0, "use_graph": True}
    with open(path) as f:
        assert json.load(f)[a]["bin_target"] == 64


def test_autotune_picks_the_fastest_candidate(monkeypatch):
    """tune() times every candidate on the caller's batch and leaves the
    context on the fastest (a fake clock makes (8x8 tiles, 64 bins, 0 per
    CU) fastest)."""
    from rusterizer_amd import autotune

    class FakeCtx:
        t = None

        def batch_tuning(self, bin_target, raster_per_cu, use_graph, fine_tile=None):
            self.t = (bin_target, raster_per_cu, use_graph, fine_tile)

    ctx = FakeCtx()
    clock = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])

    def run_batch():
        bt, rp, _, ft = ctx.t
        clock[0] += 1e-3 if (bt, rp, ft) == (64, 0, 8) else (bt + rp + ft + 1) * 1e-3

    best, results = autotune.tune(ctx, run_batch, reps=3, warmup=1)
    assert (best["bin_target"], best["raster_per_cu"], best["fine_tile"]) == (64, 0, 8)
    assert ctx.t == (64, 0, True, 8)  # left on the best
    assert len(results) == len(list(autotune.candidates()))
    assert min(results, key=lambda r: r[1])[0] == best


def test_autotune_keeps_preferred_within_tolerance(monkeypatch):
    """A preferred candidate within the tolerance of the fastest wins; one
    slower by more than the tolerance does not."""
    from rusterizer_amd import autotune

    class FakeCtx:
        t = None

        def batch_tuning(self, bin_target, raster_per_cu, use_graph, fine_tile=None):
            self.t = dict(fine_tile=fine_tile, bin_target=bin_target, raster_per_cu=raster_per_cu,
                          use_graph=use_graph)

    clock = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])
    fast = dict(fine_tile=16, bin_target=64, raster_per_cu=0, use_graph=True)
    assert fast != autotune.DEFAULT
    for pref_cost, want in ((1.015, autotune.DEFAULT), (1.05, fast)):
        ctx = FakeCtx()

        def run_batch():
            clock[0] += 1.0 if ctx.t == fast else pref_cost if ctx.t == autotune.DEFAULT else 2.0

        best, _ = autotune.tune(ctx, run_batch, reps=3, warmup=1, prefer=autotune.DEFAULT, tol=0.02)
        assert best == want and ctx.t == want
    assert autotune.DEFAULT in list(autotune.candidates())


def test_poisson_arrivals_rate_and_determinism():
    from rusterizer_amd.serve import poisson_arrivals

    a = poisson_arrivals(20000, 2000.0, seed=7)
    b = poisson_arrivals(20000, 2000.0, seed=7)
    assert np.array_equal(a, b)
    assert np.all(np.diff(a) > 0)
    gaps = np.diff(np.concatenate([[0.0], a]))
    # exponential inter-arrival times: mean 1/rate, coefficient of variation 1
    assert abs(gaps.mean() * 2000.0 - 1.0) < 0.03
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05


def test_texture_from_raw_file(tmp_path):
    from rusterizer_amd import meshes
    from rusterizer_amd.render import Texture

    tex = meshes.demo_checkerboard()
    h, w = tex.shape[:2]
    path = tmp_path / "demo.rgba"
    tex.tofile(path)
    t = Texture.from_raw_file(str(path), w, h)
    assert (t.width, t.height) == (w, h) and np.array_equal(t.buf, tex)
    # upper-left texel first, rows in order: texel (x, y) is byte 4 * (y * w + x)
    raw = np.fromfile(path, dtype=np.uint8)
    assert np.array_equal(t.buf[7, 123], raw[4 * (7 * w + 123):4 * (7 * w + 123) + 4])
    import pytest

    with pytest.raises(ValueError):
        Texture.from_raw_file(str(path), w + 1, h)
    # a 0 x N size with an empty file passes the byte count but is no texture
    empty = tmp_path / "empty.rgba"
    empty.write_bytes(b"")
    for size in ((0, 5), (5, 0), (-1, -1)):
        with pytest.raises(ValueError, match="positive"):
            Texture.from_raw_file(str(empty), *size)
