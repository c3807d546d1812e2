"""Throughput of the rasterizer on one MI355X. Prints one JSON line.

Default workload (batched): 256 independent scenes (seed 1002), each a
textured triangle soup of log-uniform 1k..200k triangles at a resolution
drawn from {256, 512, 768, 1024}^2 with its own 512x512 checkerboard, 4x
MSAA, f32 depth LESS. One step = one batch: all 256 scenes rendered and
resolved into their device framebuffers by one submission (HIP graph
replay), inputs resident in HBM. Beside it, on the same scenes:
  - the plain path: one renderer per scene, rz_draw + rz_resolve_device per
    scene, one scene per launch sequence;
  - the CPU oracle (single thread) on a bounded sample of the scenes.
The plain frames of the near-clipped textured cube (800x600, depth cleared
to 1.0), of the 1M-triangle soup (1920x1080, colour FS, seed 1001) and of
the 4M-triangle terrain (3840x2160, 4096^2 texture, seed 1003) are timed too,
each with its CPU baseline, and reported under "cube", "soup" and "terrain";
--workload cube|soup|terrain makes one of them the headline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch|soup|terrain|cube] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
PEAK_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
VALU_PER_CU_CYCLE = 2  # wave64 FP32 VALU instructions: 128 lanes per CU per clock
METRIC = "MSAA samples/s + triangles/s (batched vs plain GPU vs CPU); p99 render latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("batch", "soup", "terrain", "cube"), default="batch")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--tris", type=int, default=1_000_000, help="soup triangles")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-scene leg (batch workload)")
    ap.add_argument("--no-tune", action="store_true", help="use the default batch tuning")
    ap.add_argument("--tuning", default=None,
                    help="fixed batch tuning bin_target,raster_per_cu,use_graph[,fine_tile] (e.g. 64,0,1,16); "
                         "skips the autotuner")
    ap.add_argument("--no-extra", action="store_true", help="batch workload only: skip the soup/terrain/serve legs")
    ap.add_argument("--fine-tile", type=int, default=None,
                    help="cube/soup/terrain workloads: side of the binned screen tiles (8, 16 or 32)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson request rate (1/s)")
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--cpu-sample-tris", type=int, default=150_000,
                    help="triangles in the soup and terrain CPU samples")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy (HBM counter calibration)")
    return ap.parse_args()


def copy_probe():
    """A copy kernel of known size: reads and writes 512 MiB once. Profiling
    it with FETCH_SIZE / WRITE_SIZE gives the scale of those counters."""
    import torch

    n = 512 * 1024 * 1024 // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool((dst[:: 1 << 20] == src[:: 1 << 20]).all().item())
    print(json.dumps({"copy_probe_bytes": n * 4, "ok": ok}), flush=True)


def load_traffic(name: str):
    """HBM bytes per raster launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration); None if not measured."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            j = json.load(f)
        return j.get("bytes_per_launch"), j.get("source")
    return None, None


def log(msg: str):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_frame_seconds(sc, n_tris=None, threads: int = 1, stride: int = 1) -> float:
    """The oracle (C++ restatement of the reference; single-threaded like the
    reference unless threads > 1, which renders the same bits tile-parallel)
    on sc's first n_tris triangles (all by default), or on every stride-th
    triangle: seconds for draw + framebuffer()."""
    from oracle.oracle import OracleRenderer

    n = sc.ntri if n_tris is None else min(n_tris, sc.ntri)
    idx = sc.idx[: 3 * n]
    if stride > 1:
        idx = np.ascontiguousarray(idx.reshape(-1, 3)[::stride].reshape(-1))
    r = OracleRenderer(sc.width, sc.height, clear_depth=sc.clear_depth)
    if sc.texture is not None:
        r.bind_texture(0, sc.texture)
    r.set_uniforms(sc.world, sc.view, sc.proj)
    t0 = time.perf_counter()
    r.render(sc.pos, sc.color, sc.uv, idx, sc.fs, threads=threads)
    r.framebuffer()
    return time.perf_counter() - t0


def host_threads() -> int:
    """Host threads for the tile-parallel oracle: the box's CPU share (16)."""
    return max(1, min(16, os.cpu_count() or 1))


def raster_work(ctx, run_once, samples: int, width: int = 64, height: int = 64) -> dict:
    """Raster work counts of one more job on a context sharing ctx's meshes
    and textures, created with RZ_RASTER_STATS=1 (the counting kernel
    variant)."""
    from rusterizer_amd.render import Context

    os.environ["RZ_RASTER_STATS"] = "1"
    try:
        sctx = Context(width, height, share=ctx)
    finally:
        del os.environ["RZ_RASTER_STATS"]
    try:
        run_once(sctx)
        st = sctx.frame_stats()
    finally:
        sctx.close()
    pixels = samples // 4
    return {"list_entries": st["raster_entries"], "candidates": st["raster_candidates"],
            "prims_processed": st["raster_prims"], "covered_samples": st["raster_covered_samples"],
            "shading_tasks": st["raster_shading_tasks"],
            "covered_per_output_sample": st["raster_covered_samples"] / max(samples, 1),
            "shades_per_pixel": st["raster_shading_tasks"] / max(pixels, 1)}


def load_counters(name: str):
    """Per-launch instruction counts of the raster kernel measured with
    rocprofv3 --pmc (SQ_INSTS_VALU, ...); None if not measured."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def roofline(alg_bytes: int, raster_ms: float, traffic=(None, None), counters=None):
    """HBM roofline of the raster kernel, and beside it how much of the
    chip's instruction issue rates it uses. VALU: 157.3 TFLOP/s of FP32 FMA
    over 256 CUs at 2.4 GHz is 128 lanes per CU per clock, two wave64 VALU
    instructions per CU-cycle; SALU: one per CU-cycle (one scalar unit per
    CU). valu_issue_frac = VALU wave-instructions / (2 x CUs x clock x
    duration), salu_issue_frac likewise with 1."""
    achieved = alg_bytes / (raster_ms * 1e-3) / 1e9
    r = {"bound": "hbm", "kernel": "k_raster", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic[0], "algorithmic_bytes_per_launch": alg_bytes,
         "raster_ms": raster_ms}
    if traffic[1]:
        r["traffic_source"] = traffic[1]
    if counters:
        import torch

        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
        cycles = n_cu * PEAK_CLOCK_HZ * raster_ms * 1e-3
        r["valu_issue_frac"] = counters["SQ_INSTS_VALU"] / (VALU_PER_CU_CYCLE * cycles)
        r["salu_issue_frac"] = counters["SQ_INSTS_SALU"] / cycles
        r["issue_basis"] = (f"{counters['SQ_INSTS_VALU']:.3e} VALU and {counters['SQ_INSTS_SALU']:.3e} SALU "
                            f"wave-instructions per launch ({counters['source']}) over {n_cu} CUs x "
                            f"{PEAK_CLOCK_HZ / 1e9:.1f} GHz x {raster_ms:.3f} ms; peaks {VALU_PER_CU_CYCLE} VALU "
                            f"and 1 SALU wave-instruction per CU-cycle")
    return r


# ------------------------------------------------------------ plain frames
def bench_frame(sc, steps, warmup, cpu: str | None, traffic_name: str | None, cpu_sample_tris: int = 300_000,
                counters_name: str | None = None, fine_tile: int | None = None):
    """One scene drawn and resolved per step through the plain API.
    cpu: None, "sample" (first cpu_sample_tris triangles, time scaled --
    for i.i.d. soups), "stride" (every k-th triangle, about cpu_sample_tris
    of them, time scaled -- for ordered meshes) or "full" (the whole frame
    once); beside the 1-thread figure the whole frame is rendered
    tile-parallel on the host's threads."""
    import torch

    from rusterizer_amd.render import Context

    ctx = Context(sc.width, sc.height)
    ctx.set_clear_depth(sc.clear_depth)
    if fine_tile:
        ctx.set_fine_tile(fine_tile)
    if sc.texture is not None:
        ctx.bind_texture(0, sc.texture)
    ctx.set_uniforms(sc.world, sc.view, sc.proj)
    mesh = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
    out = torch.empty(sc.width * sc.height, dtype=torch.int32, device="cuda")
    optr = out.data_ptr()

    def frame():
        ctx.draw(mesh, sc.fs)
        ctx.resolve_device(optr)

    warm = []
    for _ in range(warmup):
        frame()
        warm.append(ctx.frame_stats())
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        f0 = time.perf_counter()
        frame()  # returns after the frame finished on the device
        lat.append(time.perf_counter() - f0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = ctx.frame_stats()
    # stage times: the same frames again with HIP events between the stages
    # (direct launches; the timed frames above replay the captured graph)
    ctx.profile(True)
    for _ in range(min(steps, PROFILED_FRAMES)):
        frame()
    torch.cuda.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    frames = prof["frames"] or min(steps, PROFILED_FRAMES)
    raster_ms = prof["raster"] / frames

    def stats_frame(c):  # the texture is already bound in the shared pool
        c.set_clear_depth(sc.clear_depth)
        c.set_uniforms(sc.world, sc.view, sc.proj)
        c.draw(mesh, sc.fs)
        c.resolve_device(optr)

    work = raster_work(ctx, stats_frame, sc.samples, sc.width, sc.height)
    res = {
        "samples_per_s": sc.samples * steps / dt,
        "triangles_per_s": sc.ntri * steps / dt,
        "ms_per_frame": dt / steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "stage_ms": {k: prof[k] / frames for k in ("geometry", "binning", "raster")},
        "frame_stats": stats,
        "raster_work": work,
        "clipped_share": stats["clipped"] / max(sc.ntri, 1),
        "warmup_replays": [(w["replays"], w["overflow_bits"]) for w in warm],
        "roofline": roofline(sc.algorithmic_bytes(), raster_ms, load_traffic(traffic_name) if traffic_name else (None, None),
                             load_counters(counters_name) if counters_name else None),
    }
    sample_kw, scale = {}, 1.0  # the CPU sample and the factor that scales its time to a frame
    if cpu == "stride":
        k = max(1, -(-sc.ntri // cpu_sample_tris))
        n = len(range(0, sc.ntri, k))
        sample_kw, scale = {"stride": k}, sc.ntri / max(n, 1)
        t = oracle_frame_seconds(sc, **sample_kw)
        frame_s = t * scale
        res["cpu_baseline"] = {
            "value": sc.samples / frame_s, "unit": "MSAA samples/s", "cores": 1, "kind": "port",
            "sample": f"every {k}th triangle ({n} of {sc.ntri}) rendered in {t:.2f} s on 1 host thread; "
                      f"frame time scaled by {sc.ntri / max(n, 1):.2f}",
            "triangles_per_s": sc.ntri / frame_s, "host_cpu": _cpu_model(), "host_nproc": os.cpu_count()}
    elif cpu == "sample":
        n = min(cpu_sample_tris, sc.ntri)
        sample_kw, scale = {"n_tris": n}, sc.ntri / max(n, 1)
        t = oracle_frame_seconds(sc, **sample_kw)
        frame_s = t * scale
        res["cpu_baseline"] = {
            "value": sc.samples / frame_s, "unit": "MSAA samples/s", "cores": 1, "kind": "port",
            "sample": f"first {n} of {sc.ntri} triangles rendered in {t:.2f} s on 1 host thread; "
                      f"frame time scaled by {sc.ntri / max(n, 1):.2f}",
            "triangles_per_s": sc.ntri / frame_s, "host_cpu": _cpu_model(), "host_nproc": os.cpu_count()}
    elif cpu == "full":
        t = oracle_frame_seconds(sc)
        res["cpu_baseline"] = {
            "value": sc.samples / t, "unit": "MSAA samples/s", "cores": 1, "kind": "port",
            "sample": f"the whole frame ({sc.ntri} triangles) rendered in {t:.2f} s on 1 host thread",
            "triangles_per_s": sc.ntri / t, "host_cpu": _cpu_model(), "host_nproc": os.cpu_count()}
    if cpu is not None:
        nt = host_threads()
        t = oracle_frame_seconds(sc, threads=nt, **sample_kw)
        res["cpu_baseline"]["multithread"] = {"value": sc.samples / (t * scale), "cores": nt, "seconds": t,
                                              "sample": "the same sample, tile-parallel (same bits)"}
        res["vs_cpu"] = res["samples_per_s"] / res["cpu_baseline"]["value"]
    mesh.destroy()
    ctx.close()
    return res


def soup_scene_of(a):
    from rusterizer_amd import scenes

    seed = a.seed if (a.seed is not None and a.workload == "soup") else scenes.SEEDS["soup"]
    sc = scenes.soup_scene(a.tris, a.width, a.height, seed=seed)
    return sc, (f"triangle soup {sc.ntri} tris, {sc.width}x{sc.height}, 4x MSAA, colour FS, "
                f"f32 depth LESS, seed {seed}")


def cube_scene_of():
    from rusterizer_amd import scenes

    sc = scenes.cube_scene(800, 600, clear_depth=1.0)
    return sc, ("cube 12 tris (mesh::cube(1.0)), 800x600, 256^2 checkerboard, 60 degree camera with the near "
                "plane cutting the cube, 4x MSAA, bilinear texture FS, f32 depth cleared to 1.0, LESS")


def terrain_scene_of(a):
    from rusterizer_amd import scenes

    seed = a.seed if (a.seed is not None and a.workload == "terrain") else scenes.SEEDS["terrain"]
    sc = scenes.terrain_scene(seed=seed)
    return sc, (f"terrain {sc.ntri} tris (1415x1415 quads, fBm heights), 3840x2160, 4096^2 checkerboard, "
                f"4x MSAA, bilinear texture FS, f32 depth LESS, seed {seed}")


# -------------------------------------------------------------------- batch
def batch_scenes_of(a):
    from rusterizer_amd import scenes

    seed = a.seed if (a.seed is not None and a.workload == "batch") else scenes.SEEDS["batch"]
    t = time.perf_counter()
    scs = scenes.batch_scenes(a.scenes, seed=seed)
    return scs, seed, time.perf_counter() - t


def bench_batch(a, steps, warmup, with_cpu, with_plain, tune, scs, seed, t_gen):
    import torch

    from rusterizer_amd import autotune, batch
    from rusterizer_amd.render import Context

    samples = sum(s.samples for s in scs)
    tris = sum(s.ntri for s in scs)
    alg_bytes = sum(s.algorithmic_bytes() for s in scs)
    ctx = Context(64, 64)
    t_up = time.perf_counter()
    items = batch.upload(ctx, scs)
    torch.cuda.synchronize()
    t_up = time.perf_counter() - t_up

    table = batch.item_table(items)  # rz_batch_item records, built once

    def once():
        batch.run(ctx, items, table)

    if a.tuning:
        v = [int(x) for x in a.tuning.split(",")]
        ctx.batch_tuning(v[0], v[1], bool(v[2]), fine_tile=v[3] if len(v) > 3 else None)
        tune = False
    once()  # first batch: buffers grow to fit
    first_stats = ctx.frame_stats()
    log(f"uploaded in {t_up:.1f} s; first batch done")
    tuning_log = None
    if tune:
        # near-equal settings (within 2%) keep the default, so that the
        # timed kernels are the ones the committed profiles describe
        best, results = autotune.tune(ctx, once, reps=2, warmup=1, prefer=autotune.DEFAULT, tol=0.02)
        tuning_log = [dict(t, ms=s * 1e3) for t, s in results]
        log(f"autotuned: {best}")
    for _ in range(warmup):
        once()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        f0 = time.perf_counter()
        once()
        lat.append(time.perf_counter() - f0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = ctx.frame_stats()
    tuning = ctx.get_batch_tuning()
    # per-stage times of the same batch launched eagerly (graph replays are
    # timed as a whole): the raster kernel's duration for the roofline
    ctx.batch_tuning(tuning["bin_target"], tuning["raster_per_cu"], False)
    once()
    ctx.profile(True)
    for _ in range(3):
        once()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.batch_tuning(**tuning)
    eager_ms = sum(prof[k] for k in ("geometry", "binning", "raster")) / prof["frames"]
    work = raster_work(ctx, lambda c: batch.run(c, items, table), samples)
    res = {
        "workload": f"batched: {len(scs)} textured soup scenes, seed {seed}, {tris} tris, resolutions "
                    f"{{256,512,768,1024}}^2 ({samples // 4} px), 512^2 checkerboard per scene, 4x MSAA",
        "samples_per_s": samples * steps / dt,
        "triangles_per_s": tris * steps / dt,
        "ms_per_batch": dt / steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "tuning": tuning,
        "autotune_ms": tuning_log,
        "stage_ms_eager": {k: prof[k] / prof["frames"] for k in ("geometry", "binning", "raster")},
        "eager_ms_per_batch": eager_ms,
        "batch_stats": stats,
        "raster_work": work,
        "first_batch_replays": (first_stats["replays"], first_stats["overflow_bits"]),
        "setup_s": {"generate": t_gen, "upload": t_up},
        "roofline": roofline(alg_bytes, prof["raster"] / prof["frames"], load_traffic("traffic_raster_batch.json"),
                             load_counters("counters_raster_batch.json")),
        "samples_per_batch": samples,
        "tris_per_batch": tris,
    }
    log(f"batched: {res['ms_per_batch']:.2f} ms per batch")
    if with_plain:
        res["plain"] = bench_plain(scs, items, steps=max(1, min(3, steps)))
        log(f"plain: {res['plain']['ms_per_pass']:.2f} ms per pass")
        res["batched_vs_plain"] = res["samples_per_s"] / res["plain"]["samples_per_s"]
    if with_cpu:
        res["cpu_baseline"] = cpu_batch(scs, a.cpu_seconds)
        log("cpu baseline done")
        res["batched_vs_cpu"] = res["samples_per_s"] / res["cpu_baseline"]["value"]
    for r in items:
        r.mesh.destroy()
    ctx.close()
    return res


SERVE_WEIGHTS = {"batch": 0.90, "soup": 0.07, "terrain": 0.03}
FRAME_REPS = 100  # timed frames of the soup and terrain legs (p50 / p99 over them)
PROFILED_FRAMES = 50  # frames run again with stage events for the stage and raster times


def bench_serve(kinds: dict, tuning: dict | None, rate: float, n: int, seed: int, slots: int = 3,
                budget: float = 200e9, service: bool = False):
    """Poisson request stream over resident templates of the three kinds,
    continuous batching over `slots` contexts (service: their rasters run in
    the group's persistent raster service); then D2H of one image of each
    kind, timed separately."""
    import torch

    from rusterizer_amd.render import Context
    from rusterizer_amd.serve import Catalog, Server, make_requests

    root = Context(64, 64)
    cat = Catalog(root, kinds, tuning=tuning)
    srv = Server(cat, slots=slots, budget_bytes=budget, max_batch=64, tuning=tuning, service=service)
    try:
        warm = make_requests(cat, min(400, n), rate, SERVE_WEIGHTS, seed + 100)
        srv.run(warm)  # buffers grow, graphs are captured
        reqs = make_requests(cat, n, rate, SERVE_WEIGHTS, seed)
        rep = srv.run(reqs, keep_outputs=False)
        d2h = {}
        for kind, ids in cat.by_kind.items():
            t = cat.templates[ids[0]]
            dev = torch.empty(t.pixels, dtype=torch.int32, device="cuda")
            host = torch.empty(t.pixels, dtype=torch.int32, pin_memory=True)
            torch.cuda.synchronize()
            ts = []
            for _ in range(10):
                t0 = time.perf_counter()
                host.copy_(dev)
                ts.append(time.perf_counter() - t0)
            ts.sort()
            d2h[kind] = {"bytes": t.pixels * 4, "ms_p50": ts[5] * 1e3, "GB_per_s": t.pixels * 4 / ts[5] / 1e9}
        rep["d2h_per_image"] = d2h
        rep["weights"] = SERVE_WEIGHTS
        rep["seed"] = seed
        rep["rate"] = rate
        return rep
    finally:
        srv.close()
        cat.destroy()
        root.close()


def bench_plain(scs, items, steps: int):
    """The same scenes through the plain API: one renderer per scene (its
    texture bound at index 0, as the reference's shader reads texture 0),
    rz_draw + rz_resolve_device per scene, one scene after the other. Checks
    that every framebuffer equals the batched one."""
    import torch

    from rusterizer_amd.render import Context

    ctxs, meshes = [], []
    for sc in scs:
        c = Context(sc.width, sc.height)
        if sc.texture is not None:
            c.bind_texture(0, sc.texture)
        c.set_clear_depth(sc.clear_depth)
        meshes.append(c.mesh(sc.pos, sc.color, sc.uv, sc.idx))
        ctxs.append(c)
    outs = [torch.empty(sc.width * sc.height, dtype=torch.int32, device="cuda") for sc in scs]

    def one_pass(lat=None):
        for sc, c, m, o in zip(scs, ctxs, meshes, outs):
            f0 = time.perf_counter()
            c.set_uniforms(sc.world, sc.view, sc.proj)
            c.draw(m, sc.fs)
            c.resolve_device(o.data_ptr())
            if lat is not None:
                lat.append(time.perf_counter() - f0)

    one_pass()  # buffers grow to fit each scene
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        one_pass(lat)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    equal = all(bool(torch.equal(o, r.out)) for o, r in zip(outs, items))
    samples = sum(s.samples for s in scs)
    tris = sum(s.ntri for s in scs)
    for m in meshes:
        m.destroy()
    for c in ctxs:
        c.close()
    return {"samples_per_s": samples * steps / dt, "triangles_per_s": tris * steps / dt,
            "ms_per_pass": dt / steps * 1e3, "passes": steps,
            "scene_latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                 "p99": float(np.percentile(lat, 99) * 1e3)},
            "images_equal_batched": equal}


def cpu_batch(scs, budget_s: float):
    """The oracle, single-threaded, on the batch's scenes in order until the
    time budget is spent (at least one whole scene); beside it the same
    scenes tile-parallel on up to 16 host threads (the same bits)."""
    done, secs, samples, tris = 0, 0.0, 0, 0
    for sc in scs:
        if secs >= budget_s and done > 0:
            break
        secs += oracle_frame_seconds(sc)
        samples += sc.samples
        tris += sc.ntri
        done += 1
    nt = host_threads()
    secs_mt = sum(oracle_frame_seconds(sc, threads=nt) for sc in scs[:done])
    return {"value": samples / secs, "unit": "MSAA samples/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of {len(scs)} scenes ({tris} tris, {samples} samples) rendered whole "
                      f"in {secs:.2f} s on 1 host thread",
            "triangles_per_s": tris / secs, "host_cpu": _cpu_model(), "host_nproc": os.cpu_count(),
            "multithread": {"value": samples / secs_mt, "cores": nt, "seconds": secs_mt}}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); multi-GPU sharding is not built",
              file=sys.stderr)
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if a.copy_probe:
        copy_probe()
        return
    with_cpu = not a.no_cpu
    if a.workload in ("soup", "terrain", "cube"):
        if a.workload == "soup":
            sc, desc = soup_scene_of(a)
            s = bench_frame(sc, a.steps, a.warmup, "sample" if with_cpu else None, "traffic_raster_soup1m.json",
                            a.cpu_sample_tris, "counters_raster_soup1m.json", a.fine_tile)
            cfg = {"workload": desc, "triangles": sc.ntri, "width": sc.width, "height": sc.height, "msaa": 4}
        elif a.workload == "terrain":
            sc, desc = terrain_scene_of(a)
            s = bench_frame(sc, a.steps, a.warmup, "full" if with_cpu else None, "traffic_raster_terrain.json",
                            a.cpu_sample_tris, "counters_raster_terrain4m.json", a.fine_tile)
            cfg = {"workload": desc, "triangles": sc.ntri, "width": sc.width, "height": sc.height, "msaa": 4,
                   "texture": 4096}
        else:
            sc, desc = cube_scene_of()
            s = bench_frame(sc, a.steps, a.warmup, "full" if with_cpu else None, None, fine_tile=a.fine_tile)
            cfg = {"workload": desc, "triangles": sc.ntri, "width": sc.width, "height": sc.height, "msaa": 4,
                   "texture": 256}
        line = {"metric": METRIC, "value": s["samples_per_s"], "unit": "MSAA samples/s", "n_gpus": 1,
                "steps": a.steps, "warmup": a.warmup, "ms_per_step": s["ms_per_frame"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": cfg,
                "triangles_per_s": s["triangles_per_s"], "latency_ms": s["latency_ms"],
                "roofline": s["roofline"], "cpu_baseline": s.get("cpu_baseline"), a.workload: s}
        print(json.dumps(line), flush=True)
        return
    scs, bseed, t_gen = batch_scenes_of(a)
    log(f"generated {len(scs)} scenes in {t_gen:.1f} s")
    b = bench_batch(a, a.steps, a.warmup, with_cpu, not a.no_plain, not a.no_tune, scs, bseed, t_gen)
    cube = soup = terrain = serve = serve_svc = None
    if a.no_extra:
        a.no_serve = True
    else:
        cpu_mode = (lambda m: m if with_cpu else None)
        cube_sc, desc = cube_scene_of()
        cube = dict(workload=desc, **bench_frame(cube_sc, FRAME_REPS, 2, cpu_mode("full"), None))
        log(f"cube: {cube['ms_per_frame']:.3f} ms per frame")
        soup_sc, desc = soup_scene_of(a)
        soup = dict(workload=desc, **bench_frame(soup_sc, FRAME_REPS, 2, cpu_mode("sample"),
                                                 "traffic_raster_soup1m.json", a.cpu_sample_tris,
                                                 "counters_raster_soup1m.json"))
        log(f"soup: {soup['ms_per_frame']:.2f} ms per frame")
        terr_sc, desc = terrain_scene_of(a)
        terrain = dict(workload=desc, **bench_frame(terr_sc, FRAME_REPS, 2, cpu_mode("full"),
                                                    "traffic_raster_terrain.json", a.cpu_sample_tris,
                                                    "counters_raster_terrain4m.json"))
        log(f"terrain: {terrain['ms_per_frame']:.2f} ms per frame")
    if not a.no_serve:
        from rusterizer_amd import scenes as S

        kinds = {"batch": scs, "soup": [soup_sc], "terrain": [terr_sc]}
        serve = bench_serve(kinds, b["tuning"], a.serve_rate, a.serve_requests, S.SEEDS["serve"])
        serve_svc = bench_serve(kinds, b["tuning"], a.serve_rate, a.serve_requests, S.SEEDS["serve"], service=True)
        for name, rp in (("serve", serve), ("serve (raster service)", serve_svc)):
            log(f"{name}: p50 {rp['latency_ms']['p50']:.2f} ms, p99 {rp['latency_ms']['p99']:.2f} ms; by kind " +
                ", ".join(f"{k} p50 {v['p50']:.2f} p99 {v['p99']:.2f}" for k, v in rp["latency_ms_by_kind"].items()))
    line = {
        "metric": METRIC, "value": b["samples_per_s"], "unit": "MSAA samples/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": b["ms_per_batch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": b["workload"], "scenes": a.scenes, "msaa": 4,
                   "samples_per_batch": b["samples_per_batch"], "triangles_per_batch": b["tris_per_batch"]},
        "triangles_per_s": b["triangles_per_s"],
        "latency_ms": b["latency_ms"],
        "serving_latency_ms": serve["latency_ms"] if serve else None,
        "serving_latency_ms_raster_service": serve_svc["latency_ms"] if serve_svc else None,
        "batched_vs_plain": b.get("batched_vs_plain"),
        "batched_vs_cpu": b.get("batched_vs_cpu"),
        "roofline": b["roofline"],
        "cpu_baseline": b.get("cpu_baseline"),
        "batch": b,
        "cube": cube,
        "soup": soup,
        "terrain": terrain,
        "serve": serve,
        "serve_raster_service": serve_svc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
