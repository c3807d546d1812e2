"""Autotuner of the batched path.

The knobs: the side of the screen tiles whose primitive lists binning
builds and the raster walks (rz_set_fine_tile: 8, 16 or 32 px); the bin
capacity, i.e. the largest number of coarse bins per scene when the bin side
is chosen automatically (the side is the smallest power of two >= 16 px
giving at most that many bins, so it sets how much screen -- and how many
primitives -- one coarse bin holds; 8x8 tiles cap the side at 128 px); the
number of persistent raster workgroups per CU (0 = one one-wave workgroup per 8x8 quadrant);
and whether the batch replays its captured HIP graph (rz_batch_tuning).
Bin lists have no fixed capacity: binning counts, scans and then writes
each list at its exact length. Every setting renders bit-identical images;
only the time changes. The tuner times each candidate on the caller's real
batch and keeps the fastest; a caller may name a preferred candidate that
wins unless another is faster by more than a tolerance, so that timing noise
between near-equal settings does not flip the choice from run to run.
"""
from __future__ import annotations

import itertools
import json
import os
import time

FINE_TILES = (8, 16, 32)
BIN_TARGETS = (16, 64, 256)
RASTER_PER_CU = (0, 8)  # 8 persistent workgroups per CU fill it (8 waves per SIMD)

# the usual winner on mixed-resolution batches of textured soups: 32x32 fine
# tiles (binning writes a quarter of the entries of 16x16 ones; the raster's
# candidate queue makes the longer lists cheap to scan), up to 16 coarse bins
# per scene, one raster workgroup per tile
DEFAULT = dict(fine_tile=32, bin_target=16, raster_per_cu=0, use_graph=True)


def candidates(fine_tiles=FINE_TILES, bin_targets=BIN_TARGETS, raster_per_cu=RASTER_PER_CU, graph=(True,)):
    for ft, bt, rp, g in itertools.product(fine_tiles, bin_targets, raster_per_cu, graph):
        yield dict(fine_tile=ft, bin_target=bt, raster_per_cu=rp, use_graph=g)


def signature(scene_sizes, tris: int) -> str:
    """Key of a batch shape: scene count, resolution histogram, triangle
    count rounded to a power of two."""
    hist = {}
    for w, h in scene_sizes:
        hist[f"{w}x{h}"] = hist.get(f"{w}x{h}", 0) + 1
    return json.dumps({"scenes": len(scene_sizes), "res": dict(sorted(hist.items())),
                       "tris_log2": max(0, int(tris).bit_length())}, sort_keys=True)


def tune(ctx, run_batch, cands=None, reps: int = 3, warmup: int = 1, prefer=None, tol: float = 0.0):
    """run_batch(): records, submits and waits one batch on ctx. Returns
    (best tuning, [(tuning, median seconds), ...]); leaves ctx on the best.
    prefer: a candidate kept as the best while its median is within
    (1 + tol) of the fastest one's."""
    results = []
    for t in (cands if cands is not None else candidates()):
        ctx.batch_tuning(**t)
        for _ in range(warmup):
            run_batch()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run_batch()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        results.append((t, ts[len(ts) // 2]))
    best, fastest = min(results, key=lambda r: r[1])
    for t, sec in results:
        if t == prefer and sec <= fastest * (1.0 + tol):
            best = t
    ctx.batch_tuning(**best)
    return best, results


class TuningCache:
    """Best tuning per batch signature, kept in a JSON file."""

    def __init__(self, path: str):
        self.path = path
        self.data = {}
        if os.path.exists(path):
            with open(path) as f:
                self.data = json.load(f)

    def get(self, sig: str):
        return self.data.get(sig)

    def put(self, sig: str, tuning: dict):
        self.data[sig] = tuning
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)
