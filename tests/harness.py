"""Shared helpers: run a Scene through the HIP library and through the oracle,
and compare them: coverage and depth bit-exact, colour (MSAA samples and the
resolved image) bit-exact too, which is tighter than the 1-LSB bar."""
from __future__ import annotations

import os

import numpy as np

CLEAR_COLOR = 0xFF191919
CPU_THREADS = max(1, min(16, os.cpu_count() or 1))


def run_gpu(scenes, coarse_bin: int | None = None, ctx=None, msaa: bool = True, fine_tile: int | None = None):
    """Render one frame made of the given scenes' draws (same size) on the GPU."""
    from rusterizer_amd.render import Context

    if not isinstance(scenes, (list, tuple)):
        scenes = [scenes]
    s0 = scenes[0]
    own = ctx is None
    if own:
        ctx = Context(s0.width, s0.height)
    try:
        ctx.set_clear_depth(s0.clear_depth)
        if coarse_bin is not None:
            ctx.set_coarse_bin(coarse_bin)
        if fine_tile is not None:
            ctx.set_fine_tile(fine_tile)
        tex_bound = False
        meshes = []
        for sc in scenes:
            if sc.texture is not None and not tex_bound:
                ctx.bind_texture(0, sc.texture)
                tex_bound = True
            ctx.set_uniforms(sc.world, sc.view, sc.proj)
            m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
            meshes.append(m)
            ctx.draw(m, sc.fs)
        if msaa:
            img, col, dep = ctx.resolve_msaa()
        else:
            img, col, dep = ctx.resolve(), None, None
        res = dict(img=img, col=col, dep=dep, flags=ctx.frame_flags(), stats=ctx.frame_stats())
        for m in meshes:
            m.destroy()
        return res
    finally:
        if own:
            ctx.close()


def run_oracle(scenes, threads: int = CPU_THREADS):
    from oracle.oracle import OracleRenderer

    if not isinstance(scenes, (list, tuple)):
        scenes = [scenes]
    s0 = scenes[0]
    r = OracleRenderer(s0.width, s0.height, clear_depth=s0.clear_depth)
    tex_bound = False
    for sc in scenes:
        if sc.texture is not None and not tex_bound:
            r.bind_texture(0, sc.texture)
            tex_bound = True
        r.set_uniforms(sc.world, sc.view, sc.proj)
        r.render(sc.pos, sc.color, sc.uv, sc.idx, sc.fs, threads=threads)
    col, dep = r.read_msaa()
    img = r.framebuffer()
    return dict(img=img, col=col, dep=dep, flags=r.flags, stats=r.stats())


def channel_diff(a: np.ndarray, b: np.ndarray) -> int:
    """Largest per-channel difference between two ARGB u32 arrays."""
    a = a.astype(np.uint32).view(np.uint8).astype(np.int16)
    b = b.astype(np.uint32).view(np.uint8).astype(np.int16)
    return int(np.abs(a - b).max()) if a.size else 0


def assert_parity(gpu: dict, cpu: dict, where: str = "", colour_lsb: int = 0):
    """Depth and coverage bit-exact; colour exact by default. The kernels
    keep the reference's operation order, no FMA contraction and the exact
    byte->float table, so colour matches the oracle bit for bit; colour_lsb=1
    relaxes it to the 1-LSB bar for a caller that documents why."""
    # depth: bit-exact per sample (this also pins coverage: every covered,
    # depth-passing sample carries its interpolated depth)
    dg = gpu["dep"].view(np.uint32)
    dc = cpu["dep"].view(np.uint32)
    bad = np.argwhere(dg != dc)
    assert bad.size == 0, f"{where}: {len(bad)} depth samples differ, first {bad[:5].tolist()}"
    # coverage: which samples hold a written colour
    wg = gpu["col"] != CLEAR_COLOR
    wc = cpu["col"] != CLEAR_COLOR
    assert np.array_equal(wg, wc), f"{where}: coverage differs at {np.argwhere(wg != wc)[:5].tolist()}"
    # colour: MSAA samples and the resolved image
    d_msaa = channel_diff(gpu["col"], cpu["col"])
    d_img = channel_diff(gpu["img"], cpu["img"])
    assert d_msaa <= colour_lsb and d_img <= colour_lsb, f"{where}: colour differs by {d_msaa} / {d_img} LSB"
    assert gpu["flags"] == cpu["flags"] & 5, f"{where}: flags gpu={gpu['flags']} cpu={cpu['flags']}"
    return dict(exact_colour=bool(np.array_equal(gpu["col"], cpu["col"]) and np.array_equal(gpu["img"], cpu["img"])))


def run_batch(scenes, ctx=None, tuning=None, msaa: bool = True, extra_draws=None, submit_only: bool = False):
    """Render scenes as one batch (rz_batch_*). scenes: list of Scene (each
    its own render target). extra_draws: {scene index: [Scene, ...]} further
    draws recorded into that scene after its own. Returns per-scene dicts
    like run_gpu (img, col, dep, flags) plus the batch stats."""
    import torch

    from rusterizer_amd.render import Context

    own = ctx is None
    if own:
        ctx = Context(64, 64)
    extra_draws = extra_draws or {}
    try:
        if tuning is not None:
            ctx.batch_tuning(**tuning)
        meshes, outs = [], []

        def prep(sc):
            t = 0
            if sc.texture is not None:
                t = ctx.texture_count()
                ctx.bind_texture(t, sc.texture)
            m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
            meshes.append(m)
            return m, t

        recs = []
        for i, sc in enumerate(scenes):
            draws = [prep(sc)] + [prep(e) for e in extra_draws.get(i, [])]
            recs.append(draws)
        ctx.batch_begin()
        for i, sc in enumerate(scenes):
            npx = sc.width * sc.height
            o = dict(img=torch.empty(npx, dtype=torch.int32, device="cuda"))
            if msaa:
                o["col"] = torch.empty(npx * 4, dtype=torch.int32, device="cuda")
                o["dep"] = torch.empty(npx * 4, dtype=torch.float32, device="cuda")
            outs.append(o)
            k = ctx.batch_scene(sc.width, sc.height, sc.clear_depth, o["img"].data_ptr())
            assert k == i
            if msaa:
                ctx.batch_scene_msaa(k, o["col"].data_ptr(), o["dep"].data_ptr())
            for (m, t), dsc in zip(recs[i], [sc] + list(extra_draws.get(i, []))):
                ctx.batch_draw(k, m, dsc.world, dsc.view, dsc.proj, t, dsc.fs)
        ctx.batch_submit()
        if submit_only:
            return ctx, outs, meshes
        ctx.batch_wait()
        res = []
        for i, (sc, o) in enumerate(zip(scenes, outs)):
            r = dict(img=o["img"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width),
                     flags=ctx.batch_scene_flags(i))
            if msaa:
                r["col"] = o["col"].cpu().numpy().view(np.uint32).reshape(sc.height, sc.width, 4)
                r["dep"] = o["dep"].cpu().numpy().reshape(sc.height, sc.width, 4)
            res.append(r)
        stats = ctx.frame_stats()
        for m in meshes:
            m.destroy()
        return res, stats
    finally:
        if own and not submit_only:
            ctx.close()
