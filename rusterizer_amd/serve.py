"""Render serving: continuous batching of requests over resident scenes.

Scene templates (meshes, textures) are uploaded once into a context group.
Requests arrive on a seeded Poisson schedule; each asks for one template to
be rendered into a framebuffer of its own. Several contexts of the group act
as batch slots, each with its own stream, so a batch can be recorded and
started while others are still running on the device. Whenever a slot is
free, the waiting requests are admitted in arrival order into one batch, as
long as the batch stays within its size limits and the device memory stays
within the HBM budget. Memory is counted in real bytes, as the library
reports them (rz_memory, rz_batch_need): the group's resident meshes and
textures, what every slot context's pipeline buffers hold (they grow to fit
the largest batch a slot has run and are kept), what the slot would hold
after the candidate batch, and the framebuffers of unfinished requests and
the pool of reusable ones. A batch that would overrun the budget is cut
down; a request that cannot fit even alone on an idle device is rejected.
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

# A fixed-rate guess of a batch's pipeline bytes per input triangle and per
# pixel; admission does not use it (it uses the library's byte counts), it
# is kept to compare against them.
BYTES_PER_TRI = 1200
BYTES_PER_PX = 64


@dataclasses.dataclass
class Template:
    kind: str        # "soup" | "batch" | "terrain"
    scene: object    # scenes.Scene
    mesh: object     # DeviceMesh
    tex: int
    row: np.ndarray  # rz_batch_item record (argb_dev filled per request)
    need: int = 0    # pipeline bytes it took, rendered alone on an empty context (beyond that context's own)

    @property
    def pixels(self) -> int:
        return self.scene.width * self.scene.height


class Catalog:
    """Templates resident in the group of `root`. Each template is rendered
    once alone on an emptied context of the group, which measures the
    pipeline bytes it really takes, bin lists grown to its real triangle
    sizes included (the library's up-front estimate knows only triangle and
    pixel counts)."""

    def __init__(self, root: Context, scenes_by_kind: dict, tuning: dict | None = None):
        """tuning: the slots' rz_batch_tuning (the measurement uses it too)."""
        self.root = root
        self.tuning = dict(tuning or {})
        self.templates: list[Template] = []
        self.by_kind: dict[str, list[int]] = collections.defaultdict(list)
        for kind, scs in scenes_by_kind.items():
            for sc in scs:
                t = 0
                if sc.texture is not None:
                    t = root.texture_count()
                    root.bind_texture(t, sc.texture)
                m = root.mesh(sc.pos, sc.color, sc.uv, sc.idx)
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
        self._measure()

    def _measure(self):
        import torch

        probe = Context(64, 64, share=self.root)
        try:
            probe.batch_tuning(**dict(self.tuning, use_graph=False))
            probe.trim()
            self.empty_bytes = probe.memory()["held"]
            for t in self.templates:
                img = torch.empty(t.pixels, dtype=torch.int32, device="cuda")
                row = np.array([t.row], dtype=L.batch_item_dtype())
                row[0]["argb_dev"] = img.data_ptr()
                probe.batch_begin()
                probe.batch_add(row)
                probe.batch_submit()
                probe.batch_wait()
                t.need = probe.memory()["held"] - self.empty_bytes
                probe.trim()
                del img
        finally:
            probe.close()

    @property
    def resident_bytes(self) -> int:
        """Device bytes of the group's meshes and textures (rz_memory)."""
        return self.root.memory()["resident"]

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
    """Reusable device framebuffers, by pixel count; counts the bytes of the
    framebuffers it has handed out and of those it keeps for reuse."""

    def __init__(self):
        self.free = collections.defaultdict(list)
        self.cached_bytes = 0
        self.out_bytes = 0

    def get(self, npx: int):
        import torch

        lst = self.free[npx]
        self.out_bytes += 4 * npx
        if lst:
            self.cached_bytes -= 4 * npx
            return lst.pop()
        return torch.empty(npx, dtype=torch.int32, device="cuda")

    def new_bytes(self, pixels) -> int:
        """Device bytes that handing out framebuffers of these pixel counts
        would allocate (beyond the cached ones)."""
        need = collections.Counter(pixels)
        return sum(4 * n * max(0, k - len(self.free[n])) for n, k in need.items())

    def put(self, t):
        n = int(t.numel())
        self.out_bytes -= 4 * n
        self.cached_bytes += 4 * n
        self.free[n].append(t)


class Server:
    """Continuous batching over `slots` contexts of the catalog's group."""

    def __init__(self, catalog: Catalog, slots: int = 3, budget_bytes: float = 200e9, max_batch: int = 64,
                 max_batch_tris: int = 12_000_000, tuning: dict | None = None, small_cus: int = 0,
                 small_tris: int = 250_000, small_px: int = 1 << 20, service: bool = False,
                 service_wg_per_cu: int = 2, service_idle_us: int = 50_000):
        """small_cus > 0 (with slots >= 2): slot 0 serves only small requests
        (at most small_tris triangles and small_px pixels) on CUs
        [0, small_cus); the other slots serve any request on the remaining
        CUs, so a small request never waits behind a long job's workgroups.
        service: the slots post their batches' raster to the group's
        persistent raster service (rz_service_*), which renders the tiles of
        all slots' batches in one resident kernel, the batch with the fewest
        tiles left first."""
        self.catalog = catalog
        self.ctxs = [Context(64, 64, share=catalog.root) for _ in range(slots)]
        for c in self.ctxs:
            if tuning:
                c.batch_tuning(**tuning)
        self.service = service
        if service:
            catalog.root.service_start(service_wg_per_cu, service_idle_us)
            for c in self.ctxs:
                c.service_attach(True)
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
        if self.service:
            self.catalog.root.service_stop()
            self.service = False

    def is_small(self, tmpl: Template) -> bool:
        return tmpl.scene.ntri <= self.small_tris and tmpl.pixels <= self.small_px

    def est_bytes(self, tmpl: Template) -> int:
        """The fixed-rate guess of one request's device bytes (not used for
        admission; compared against the library's counts)."""
        return tmpl.pixels * (4 + BYTES_PER_PX) + tmpl.scene.ntri * BYTES_PER_TRI

    def device_bytes(self, resident: int, held: list[int]) -> int:
        """Device bytes in use: resident data, every slot's buffers (as given),
        framebuffers handed out (kept outputs included) and kept for reuse."""
        return resident + sum(held) + self.pool.out_bytes + self.pool.cached_bytes

    def run(self, reqs: list[Request], copy_back: bool = False, keep_outputs: bool = False,
            timeout_s: float = 120.0) -> dict:
        """Serve reqs (sorted by arrival). keep_outputs: leave each request's
        framebuffer in req.out (else it returns to the pool)."""
        cat = self.catalog.templates
        # waiting requests in arrival order; with a small-request slot the small
        # ones queue apart, so that slot never scans past large requests
        small_q: collections.deque[Request] = collections.deque()
        large_q: collections.deque[Request] = collections.deque()
        busy: list[list[Request] | None] = [None] * len(self.ctxs)
        rm = self.catalog.root.memory()
        resident = rm["resident"] + rm["held"]  # the group's data and the root context's own buffers
        held = [c.memory()["held"] for c in self.ctxs]
        trimmed_for = None  # the request the idle slots last gave memory back for
        trims = 0
        peak_bytes = self.device_bytes(resident, held)
        nxt = 0
        done = 0
        batches = []
        cuts = 0
        overlap_polls = busy_polls = 0
        max_concurrent = 0

        def head(small_only: bool):
            """The oldest waiting request this slot may take, and its queue."""
            if small_only:
                return (small_q[0], small_q) if small_q else (None, None)
            a = small_q[0] if small_q else None
            b = large_q[0] if large_q else None
            if a is None or (b is not None and b.t_arrival < a.t_arrival):
                return (b, large_q) if b is not None else (None, None)
            return a, small_q

        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            if now > timeout_s:
                raise TimeoutError(f"serving did not finish in {timeout_s} s ({done}/{len(reqs)} done)")
            while nxt < len(reqs) and reqs[nxt].t_arrival <= now:
                r = reqs[nxt]
                (small_q if self.small_slot and self.is_small(cat[r.tmpl]) else large_q).append(r)
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
                        if copy_back:
                            c0 = time.perf_counter()
                            r.image = r.out.cpu()
                            r.d2h_s = time.perf_counter() - c0
                        if not keep_outputs:
                            self.pool.put(r.out)
                            r.out = None
                    done += len(busy[s])
                    busy[s] = None
                    held[s] = ctx.memory()["held"]  # grown further if the batch outgrew its estimate
                    peak_bytes = max(peak_bytes, self.device_bytes(resident, held))
                else:
                    nbusy += 1
            # admission: fill every free slot from the queues, oldest first
            blocked_r, blocked = None, set()  # a request that fits no slot alone, and where it failed
            for s, ctx in enumerate(self.ctxs):
                if busy[s] is not None or not (small_q or large_q):
                    continue
                small_only = self.small_slot and s == 0
                take: list[tuple[Request, collections.deque]] = []
                tris = 0
                while len(take) < self.max_batch:
                    r, q = head(small_only)
                    if r is None:
                        break
                    t = cat[r.tmpl]
                    if take and tris + t.scene.ntri > self.max_batch_tris:
                        break
                    q.popleft()
                    tris += t.scene.ntri
                    take.append((r, q))
                if not take:
                    continue
                # cut the batch until the device bytes it leads to fit the budget.
                # The slot's bytes after the batch: the library's estimate from
                # the triangle and pixel counts (rz_batch_need), or, if larger,
                # the sum of what the batch's templates took alone (their bin
                # lists at their real lengths) -- buffers never shrink, so at
                # least what the slot holds already.
                while True:
                    rows = np.empty(len(take), dtype=L.batch_item_dtype())
                    for i, (r, _) in enumerate(take):
                        rows[i] = cat[r.tmpl].row
                    measured = self.catalog.empty_bytes + sum(cat[r.tmpl].need for r, _ in take)
                    need = max(ctx.batch_need(rows), measured, held[s])
                    fb_new = self.pool.new_bytes([cat[r.tmpl].pixels for r, _ in take])
                    total = self.device_bytes(resident, held[:s] + [need] + held[s + 1:]) + fb_new
                    if total <= self.budget or len(take) == 1:
                        break
                    keep = len(take) // 2
                    for r, q in reversed(take[keep:]):
                        q.appendleft(r)
                    take = take[:keep]
                    cuts += 1
                if total > self.budget:
                    r, q = take[0]
                    q.appendleft(r)  # wait for memory to free up
                    if blocked_r is not r:
                        blocked_r, blocked = r, set()
                    blocked.add(s)
                    continue
                for i, (r, _) in enumerate(take):
                    r.out = self.pool.get(cat[r.tmpl].pixels)
                    rows[i]["argb_dev"] = r.out.data_ptr()
                    r.nbytes = need
                ctx.batch_begin()
                ctx.batch_add(rows)
                ctx.batch_submit()
                held[s] = need
                peak_bytes = max(peak_bytes, self.device_bytes(resident, held))
                ts = time.perf_counter() - t0
                for r, _ in take:
                    r.t_submit = ts
                    r.slot = s
                    r.batch = len(batches)
                batches.append(len(take))
                busy[s] = [r for r, _ in take]
                nbusy += 1
            if blocked_r is not None and nbusy == 0:
                # nothing runs, so nothing frees memory: the idle slots give
                # their buffers back once; a request that still fits on no
                # slot that may take it never will
                small = self.small_slot and self.is_small(cat[blocked_r.tmpl])
                eligible = set(range(len(self.ctxs))) - ({0} if self.small_slot and not small else set())
                if eligible <= blocked:
                    if trimmed_for is not blocked_r:
                        for s2, c2 in enumerate(self.ctxs):
                            c2.trim()
                            held[s2] = c2.memory()["held"]
                        trimmed_for = blocked_r
                        trims += 1
                    else:
                        q = small_q if small else large_q
                        q.remove(blocked_r)
                        blocked_r.rejected = True
                        done += 1
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
            "service": self.service,
            "max_batches_in_flight": max_concurrent,
            "overlap_share": overlap_polls / max(busy_polls, 1),
            "budget_bytes": self.budget,
            "peak_device_bytes": peak_bytes,
            "resident_bytes": self.catalog.resident_bytes,
            "slot_bytes": [c.memory()["held"] for c in self.ctxs],
            "budget_cuts": cuts,
            "trims": trims,
            "shared_batch_share": sum(b for b in batches if b >= 2) / max(sum(batches), 1),
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
