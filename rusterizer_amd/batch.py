"""Resident scene sets for the batched path: meshes, textures and
framebuffers uploaded once and kept on the device, recorded into batches
(rz_batch_*) on any context of the same group."""
from __future__ import annotations

import numpy as np

from .render import Context, DeviceMesh


class Resident:
    """One scene kept on the device: its mesh, texture index and framebuffer."""

    def __init__(self, scene, mesh: DeviceMesh, tex: int, out):
        self.scene = scene
        self.mesh = mesh
        self.tex = tex
        self.out = out  # torch int32 tensor, width*height words (ARGB)

    @property
    def samples(self) -> int:
        return 4 * self.scene.width * self.scene.height


def upload(ctx: Context, scenes, torch_device: str = "cuda") -> list[Resident]:
    """Upload every scene's mesh and texture into ctx's group and allocate
    its framebuffer."""
    import torch

    out = []
    for sc in scenes:
        t = 0
        if sc.texture is not None:
            t = ctx.texture_count()
            ctx.bind_texture(t, sc.texture)
        m = ctx.mesh(sc.pos, sc.color, sc.uv, sc.idx)
        fb = torch.empty(sc.width * sc.height, dtype=torch.int32, device=torch_device)
        out.append(Resident(sc, m, t, fb))
    return out


def item_table(items: list[Resident]) -> np.ndarray:
    """rz_batch_item records of the items (one scene, one draw each)."""
    from . import _lib as L

    t = np.zeros(len(items), dtype=L.batch_item_dtype())
    for i, r in enumerate(items):
        sc = r.scene
        t[i]["width"], t[i]["height"] = sc.width, sc.height
        t[i]["clear_depth"] = sc.clear_depth
        t[i]["texture"] = r.tex
        t[i]["vertex_shader"] = L.RZ_VS_MVP
        t[i]["fragment_shader"] = sc.fs
        t[i]["argb_dev"] = r.out.data_ptr()
        t[i]["mesh"] = r.mesh.handle.value
        t[i]["world"] = np.asarray(sc.world, np.float32).reshape(16)
        t[i]["view"] = np.asarray(sc.view, np.float32).reshape(16)
        t[i]["proj"] = np.asarray(sc.proj, np.float32).reshape(16)
    return t


def record(ctx: Context, items: list[Resident], table: np.ndarray | None = None):
    """Record one batch drawing each item into its own framebuffer (one
    rz_batch_add call; pass a prebuilt item_table to skip rebuilding it)."""
    ctx.batch_begin()
    ctx.batch_add(item_table(items) if table is None else table)


def run(ctx: Context, items: list[Resident], table: np.ndarray | None = None):
    record(ctx, items, table)
    ctx.batch_submit()
    ctx.batch_wait()


def image(r: Resident) -> np.ndarray:
    return r.out.cpu().numpy().view(np.uint32).reshape(r.scene.height, r.scene.width)
