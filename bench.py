"""Throughput of the rasterizer hot path on one MI355X.

Workload (default): the synthetic triangle soup -- 1,000,000 triangles,
1920x1080, 4x MSAA, colour fragment shader, f32 depth LESS, seed 1001.
One step = one frame: rz_draw of the device-resident mesh + rz_resolve_device
(vertex stage, clipping, setup, binning, raster, resolve) with inputs already
in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample-tris", type=int, default=300_000)
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


def load_traffic():
    """HBM bytes per raster launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy-kernel calibration); None if not measured."""
    p = os.path.join(ROOT, "profiles", "traffic_raster_soup1m.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def cpu_baseline(sc, n_sample: int):
    """The oracle (single-threaded C++ restatement of the reference) on the
    first n_sample triangles of the same soup (about 10 s of CPU work); the
    frame time is scaled by n_tris / n_sample (triangles are i.i.d., so the
    sample is representative)."""
    from oracle.oracle import OracleRenderer

    n_sample = min(n_sample, sc.ntri)
    idx = sc.idx[: 3 * n_sample]
    r = OracleRenderer(sc.width, sc.height, clear_depth=sc.clear_depth)
    r.set_uniforms(sc.world, sc.view, sc.proj)
    t0 = time.perf_counter()
    r.render(sc.pos, sc.color, sc.uv, idx, sc.fs, threads=1)
    r.framebuffer()
    dt = time.perf_counter() - t0
    frame_s = dt * sc.ntri / max(n_sample, 1)
    return {
        "value": sc.samples / frame_s,
        "unit": "MSAA samples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first {n_sample} of {sc.ntri} triangles rendered in {dt:.2f} s on 1 host thread; "
                  f"frame time scaled by {sc.ntri / max(n_sample, 1):.2f}",
        "triangles_per_s": sc.ntri / frame_s,
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
    from rusterizer_amd import scenes
    from rusterizer_amd.render import Context

    sc = scenes.soup_scene(a.tris, a.width, a.height, seed=a.seed)
    ctx = Context(sc.width, sc.height)
    ctx.set_uniforms(sc.world, sc.view, sc.proj)
    mesh = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
    out = torch.empty(sc.width * sc.height, dtype=torch.int32, device="cuda")
    optr = out.data_ptr()

    def frame():
        ctx.draw(mesh, sc.fs)
        ctx.resolve_device(optr)

    warm_stats = []
    for _ in range(a.warmup):
        frame()
        warm_stats.append(ctx.frame_stats())
    torch.cuda.synchronize()
    ctx.profile(True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        f0 = time.perf_counter()
        frame()  # returns after the frame finished on the device
        lat.append(time.perf_counter() - f0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = ctx.profile_read()
    stats = ctx.frame_stats()
    ctx.profile(False)

    frames = prof["frames"] or a.steps
    raster_ms = prof["raster"] / frames
    alg_bytes = sc.algorithmic_bytes()
    achieved = alg_bytes / (raster_ms * 1e-3) / 1e9
    ms_per_step = dt / a.steps * 1e3
    value = sc.samples * a.steps / dt
    line = {
        "metric": "MSAA samples/s",
        "value": value,
        "unit": "MSAA samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"triangle soup {sc.ntri} tris, {sc.width}x{sc.height}, 4x MSAA, colour FS, "
                        f"f32 depth LESS, seed {a.seed}",
            "triangles": sc.ntri,
            "width": sc.width,
            "height": sc.height,
            "msaa": 4,
        },
        "triangles_per_s": sc.ntri * a.steps / dt,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "stage_ms": {k: prof[k] / frames for k in ("geometry", "binning", "raster")},
        "frame_stats": stats,
        "warmup_replays": [(w["replays"], w["overflow_bits"]) for w in warm_stats],
        "roofline": {
            "bound": "hbm",
            "kernel": "k_raster",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, a.cpu_sample_tris)
    else:
        line["cpu_baseline"] = None
    mesh.destroy()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
