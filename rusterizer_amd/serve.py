"""Render serving: continuous batching of requests over resident scenes.

Scene templates (meshes, textures) are uploaded once into a context group.
Requests arrive on a seeded Poisson schedule; each asks for one template to
be rendered into a framebuffer of its own. Several contexts of the group act
as batch slots, each with its own stream, so a batch can be recorded and
started while others are still running on the device. Whenever a slot is
free, the waiting requests are admitted in arrival order into one batch, as
long as the batch stays within its size limits and the estimated device
memory of all admitted, unfinished requests stays within the HBM budget.
A request's latency runs from its arrival to the completion of its batch
on the device (observed by polling); copying the image back to the host is
timed separately.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _lib as L
from .render import Context

# device bytes per input triangle and per pixel that a batch's pipeline
# buffers take (prim records, attributes, bin lists), used for admission
BYTES_PER_TRI = 1200
BYTES_PER_PX = 64


@dataclasses.dataclass
class Template:
    kind: str        # "soup" | "batch" | "terrain"
    scene: object    # scenes.Scene
    mesh: object     # DeviceMesh
    tex: int
    row: np.ndarray  # rz_batch_item record (argb_dev filled per request)

    @property
    def pixels(self) -> int:
        return self.scene.width * self.scene.height


class Catalog:
    """Templates resident in the group of `root`."""

    def __init__(self, root: Context, scenes_by_kind: dict):
        self.root = root
        self.templates: list[Template] = []
        self.by_kind: dict[str, list[int]] = collections.defaultdict(list)
        self.resident_bytes = 0
        for kind, scs in scenes_by_kind.items():
            for sc in scs:
                t = 0
                if sc.texture is not None:
                    t = root.texture_count()
                    root.bind_texture(t, sc.texture)
                    self.resident_bytes += sc.texture.nbytes
                m = root.mesh(sc.pos, sc.color, sc.uv, sc.idx)
                self.resident_bytes += sc.pos.nbytes + sc.idx.nbytes
                self.resident_bytes += sc.color.nbytes if sc.color is not None else 0
                self.resident_bytes += sc.uv.nbytes if sc.uv is not None else 0
                row = np.zeros(1, dtype=L.batch_item_dtype())[0]
                row["width"], row["height"] = sc.width, sc.height
                row["clear_depth"] = sc.clear_depth
                row["texture"] = t
                row["vertex_shader"] = L.RZ_VS_MVP
                row["fragment_shader"] = sc.fs
                row["mesh"] = m.handle.value
                row["world"] = np.asarray(sc.world, np.float32).reshape(16)
                row["view"] = np.asarray(sc.view, np.float32).reshape(16)
                row["proj"] = np.asarray(sc.proj, np.float32).reshape(16)
                self.by_kind[kind].append(len(self.templates))
                self.templates.append(Template(kind, sc, m, t, row))

    def destroy(self):
        for t in self.templates:
            t.mesh.destroy()
        self.templates = []


@dataclasses.dataclass
class Request:
    rid: int
    tmpl: int
    t_arrival: float
    t_submit: float = -1.0
    t_done: float = -1.0
    d2h_s: float = 0.0
    out: object = None  # torch int32 device tensor
    image: object = None  # host copy (copy_back)
    rejected: bool = False  # larger than the whole budget on its own
    nbytes: int = 0
    slot: int = -1
    batch: int = -1


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """n arrival times (s) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def make_requests(catalog: Catalog, n: int, rate: float, weights: dict, seed: int) -> list[Request]:
    """Requests on a Poisson schedule; each picks a kind with the given
    weights, then a template of that kind uniformly."""
    rng = np.random.default_rng(seed + 1)
    kinds = [k for k in weights if catalog.by_kind.get(k)]
    p = np.array([weights[k] for k in kinds], float)
    p /= p.sum()
    arr = poisson_arrivals(n, rate, seed)
    out = []
    for i in range(n):
        k = kinds[int(rng.choice(len(kinds), p=p))]
        ids = catalog.by_kind[k]
        out.append(Request(i, ids[int(rng.integers(0, len(ids)))], float(arr[i])))
    return out


class FramePool:
    """Reusable device framebuffers, by pixel count."""

    def __init__(self):
        self.free = collections.defaultdict(list)

    def get(self, npx: int):
        import torch

        lst = self.free[npx]
        return lst.pop() if lst else torch.empty(npx, dtype=torch.int32, device="cuda")

    def put(self, t):
        self.free[int(t.numel())].append(t)


class Server:
    """Continuous batching over `slots` contexts of the catalog's group."""

    def __init__(self, catalog: Catalog, slots: int = 3, budget_bytes: float = 200e9, max_batch: int = 64,
                 max_batch_tris: int = 12_000_000, tuning: dict | None = None, small_cus: int = 0,
                 small_tris: int = 250_000, small_px: int = 1 << 20):
        """small_cus > 0 (with slots >= 2): slot 0 serves only small requests
        (at most small_tris triangles and small_px pixels) on CUs
        [0, small_cus); the other slots serve any request on the remaining
        CUs, so a small request never waits behind a long job's workgroups."""
        self.catalog = catalog
        self.ctxs = [Context(64, 64, share=catalog.root) for _ in range(slots)]
        for c in self.ctxs:
            if tuning:
                c.batch_tuning(**tuning)
        self.small_slot = small_cus > 0 and slots >= 2
        self.small_tris, self.small_px = small_tris, small_px
        if self.small_slot:
            import torch

            n_cu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
            if not 0 < small_cus < n_cu:
                raise ValueError(f"small_cus must be in (0, {n_cu})")
            self.ctxs[0].set_cu_mask(range(small_cus))
            for c in self.ctxs[1:]:
                c.set_cu_mask(range(small_cus, n_cu))
        self.budget = budget_bytes
        self.max_batch = max_batch
        self.max_batch_tris = max_batch_tris
        self.pool = FramePool()

    def close(self):
        for c in self.ctxs:
            c.close()
        self.ctxs = []

    def is_small(self, tmpl: Template) -> bool:
        return tmpl.scene.ntri <= self.small_tris and tmpl.pixels <= self.small_px

    def est_bytes(self, tmpl: Template) -> int:
        return tmpl.pixels * (4 + BYTES_PER_PX) + tmpl.scene.ntri * BYTES_PER_TRI

    def run(self, reqs: list[Request], copy_back: bool = False, keep_outputs: bool = False,
            timeout_s: float = 120.0) -> dict:
        """Serve reqs (sorted by arrival). keep_outputs: leave each request's
        framebuffer in req.out (else it returns to the pool)."""
        cat = self.catalog.templates
        queue: collections.deque[Request] = collections.deque()
        busy: list[list[Request] | None] = [None] * len(self.ctxs)
        in_flight_bytes = self.catalog.resident_bytes
        peak_bytes = in_flight_bytes
        nxt = 0
        done = 0
        batches = []
        overlap_polls = busy_polls = 0
        max_concurrent = 0
        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            if now > timeout_s:
                raise TimeoutError(f"serving did not finish in {timeout_s} s ({done}/{len(reqs)} done)")
            while nxt < len(reqs) and reqs[nxt].t_arrival <= now:
                queue.append(reqs[nxt])
                nxt += 1
            # completions
            nbusy = 0
            for s, ctx in enumerate(self.ctxs):
                if busy[s] is None:
                    continue
                if ctx.batch_query():
                    tdone = time.perf_counter() - t0
                    for r in busy[s]:
                        r.t_done = tdone
                        in_flight_bytes -= r.nbytes
                        if copy_back:
                            c0 = time.perf_counter()
                            r.image = r.out.cpu()
                            r.d2h_s = time.perf_counter() - c0
                        if not keep_outputs:
                            self.pool.put(r.out)
                            r.out = None
                    done += len(busy[s])
                    busy[s] = None
                else:
                    nbusy += 1
            # admission: fill every free slot from the queue, oldest first (the
            # small-request slot passes over large requests, leaving them queued)
            for s, ctx in enumerate(self.ctxs):
                if busy[s] is not None or not queue:
                    continue
                small_only = self.small_slot and s == 0
                take: list[Request] = []
                skipped: list[Request] = []
                tris = 0
                while queue and len(take) < self.max_batch:
                    r = queue[0]
                    t = cat[r.tmpl]
                    need = self.est_bytes(t)
                    if self.catalog.resident_bytes + need > self.budget:
                        queue.popleft()  # can never fit: reject
                        r.rejected = True
                        done += 1
                        continue
                    if small_only and not self.is_small(t):
                        skipped.append(queue.popleft())
                        continue
                    if take and tris + t.scene.ntri > self.max_batch_tris:
                        break
                    if in_flight_bytes + need > self.budget:
                        break  # wait for memory to free up
                    queue.popleft()
                    r.nbytes = need
                    in_flight_bytes += need
                    tris += t.scene.ntri
                    take.append(r)
                queue.extendleft(reversed(skipped))
                if not take:
                    continue
                peak_bytes = max(peak_bytes, in_flight_bytes)
                rows = np.empty(len(take), dtype=L.batch_item_dtype())
                for i, r in enumerate(take):
                    t = cat[r.tmpl]
                    r.out = self.pool.get(t.pixels)
                    rows[i] = t.row
                    rows[i]["argb_dev"] = r.out.data_ptr()
                ctx.batch_begin()
                ctx.batch_add(rows)
                ctx.batch_submit()
                ts = time.perf_counter() - t0
                for r in take:
                    r.t_submit = ts
                    r.slot = s
                    r.batch = len(batches)
                batches.append(len(take))
                busy[s] = take
                nbusy += 1
            if nbusy:
                busy_polls += 1
                if nbusy >= 2:
                    overlap_polls += 1
                max_concurrent = max(max_concurrent, nbusy)
        makespan = time.perf_counter() - t0
        served = [r for r in reqs if not r.rejected]
        if not served:
            return {"requests": 0, "rejected": len(reqs), "makespan_s": makespan, "batches": len(batches)}
        lat = np.array([r.t_done - r.t_arrival for r in served])
        queue_wait = np.array([r.t_submit - r.t_arrival for r in served])
        samples = sum(4 * cat[r.tmpl].pixels for r in served)
        tris = sum(cat[r.tmpl].scene.ntri for r in served)
        kinds = collections.Counter(cat[r.tmpl].kind for r in served)
        rep = {
            "requests": len(served),
            "rejected": len(reqs) - len(served),
            "kinds": dict(kinds),
            "offered_rate": len(reqs) / max(reqs[-1].t_arrival, 1e-9),
            "makespan_s": makespan,
            "requests_per_s": len(served) / makespan,
            "samples_per_s": samples / makespan,
            "triangles_per_s": tris / makespan,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                           "mean": float(lat.mean() * 1e3), "max": float(lat.max() * 1e3)},
            "queue_ms": {"p50": float(np.percentile(queue_wait, 50) * 1e3),
                         "p99": float(np.percentile(queue_wait, 99) * 1e3)},
            "batches": len(batches),
            "batch_size": {"mean": float(np.mean(batches)), "max": int(max(batches))},
            "slots": len(self.ctxs),
            "small_slot": self.small_slot,
            "max_batches_in_flight": max_concurrent,
            "overlap_share": overlap_polls / max(busy_polls, 1),
            "budget_bytes": self.budget,
            "peak_admitted_bytes": peak_bytes,
            "resident_bytes": self.catalog.resident_bytes,
        }
        by_kind = collections.defaultdict(list)
        for r, l in zip(served, lat):
            by_kind[cat[r.tmpl].kind].append(l)
        rep["latency_ms_by_kind"] = {
            k: {"n": len(v), "p50": float(np.percentile(v, 50) * 1e3), "p99": float(np.percentile(v, 99) * 1e3)}
            for k, v in sorted(by_kind.items())}
        if copy_back:
            d2h = np.array([r.d2h_s for r in served])
            nbytes = sum(4 * cat[r.tmpl].pixels for r in served)
            rep["d2h"] = {"total_s": float(d2h.sum()), "bytes": nbytes, "GB_per_s": nbytes / max(d2h.sum(), 1e-9) / 1e9,
                          "p50_ms": float(np.percentile(d2h, 50) * 1e3)}
        return rep
