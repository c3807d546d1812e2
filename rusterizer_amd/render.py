"""Host-side mirror of rusterizer's draw API over the HIP library.

`Renderer`, `Uniforms`, `UniformBlock`, `Mesh`, `Texture` keep the names,
argument meaning and call order of src/render.rs, src/uniform.rs,
src/mesh.rs and src/texture.rs; shaders, which are Rust fn pointers there,
are enum values here (`VertexShader.MVP`, `FragmentShader.{TEXTURE,COLOR,DEBUG}`,
the closures of src/main.rs:67-77 and :147-152) because the code runs on
the GPU. `Context` is the thin wrapper of the C ABI that Renderer uses.
"""
from __future__ import annotations

import ctypes as C
import enum
import weakref

import numpy as np

from . import _lib as L


class VertexShader(enum.IntEnum):
    MVP = L.RZ_VS_MVP  # main.rs:147-152


class FragmentShader(enum.IntEnum):
    TEXTURE = L.RZ_FS_TEXTURE  # main.rs:69-71
    COLOR = L.RZ_FS_COLOR      # main.rs:72
    DEBUG = L.RZ_FS_DEBUG      # main.rs:73-75


class RzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rz error {code}: {msg}")
        self.code = code


def _f32(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape_last is not None:
        a = a.reshape(-1, shape_last)
    return a


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class DeviceMesh:
    """A mesh resident on the device (rz_mesh)."""

    def __init__(self, ctx: "Context", handle, ntri: int):
        self._ctx = ctx
        self.handle = handle
        self.ntri = ntri

    def destroy(self):
        if self.handle:
            self._ctx._check(L.load().rz_mesh_destroy(self._ctx.handle, self.handle))
            self.handle = None


class Context:
    """One rz_ctx: a headless renderer with its own HIP stream."""

    def __init__(self, width: int, height: int, share: "Context | None" = None):
        """share: an existing context whose meshes and textures this one uses
        (rz_create_shared); its own stream and buffers are separate."""
        self._L = L.load()
        h = C.c_void_p()
        if share is None:
            rc = self._L.rz_create(width, height, C.byref(h))
        else:
            rc = self._L.rz_create_shared(share.handle, width, height, C.byref(h))
        if rc != 0:
            raise RzError(rc, "rz_create failed")
        self.handle = h
        self._share = share  # keeps the group's first context alive no longer than needed
        self.width, self.height = width, height

    def close(self):
        if getattr(self, "handle", None):
            self._L.rz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            msg = self._L.rz_last_error(self.handle)
            raise RzError(rc, msg.decode() if msg else "")

    # ---- state
    @property
    def stream(self) -> int:
        return int(self._L.rz_stream(self.handle) or 0)

    def set_cu_mask(self, cus=None):
        """Restrict this context's work to the given CU indices (None: all
        CUs); the context must be idle. Replaces its stream."""
        if cus is None:
            self._check(self._L.rz_set_cu_mask(self.handle, None, 0))
            return
        cus = [int(i) for i in cus]
        words = (max(cus) // 32 + 1) if cus else 1
        m = (C.c_uint32 * words)()
        for i in cus:
            m[i // 32] |= 1 << (i % 32)
        self._check(self._L.rz_set_cu_mask(self.handle, m, words))

    def set_clear_depth(self, depth: float):
        self._check(self._L.rz_set_clear_depth(self.handle, float(depth)))

    def set_uniforms(self, world, view, proj):
        w, v, p = (_f32(m).reshape(16) for m in (world, view, proj))
        fp = C.POINTER(C.c_float)
        self._check(self._L.rz_set_uniforms(self.handle, w.ctypes.data_as(fp), v.ctypes.data_as(fp),
                                            p.ctypes.data_as(fp)))

    def bind_texture(self, index: int, rgba8: np.ndarray):
        t = np.ascontiguousarray(rgba8, dtype=np.uint8)
        if t.ndim != 3 or t.shape[2] != 4:
            raise ValueError("texture must be (height, width, 4) RGBA8")
        th, tw = t.shape[:2]
        self._check(self._L.rz_bind_texture(self.handle, index, t.ctypes.data_as(C.POINTER(C.c_uint8)), tw, th))

    def texture_count(self) -> int:
        return int(self._L.rz_texture_count(self.handle))

    def memory(self) -> dict:
        """rz_memory: device bytes held by this context, resident in its
        group, and what it will hold once the recorded job has run."""
        a = (C.c_uint64 * 3)()
        self._check(self._L.rz_memory(self.handle, a))
        return dict(held=int(a[0]), resident=int(a[1]), job=int(a[2]))

    def batch_need(self, items) -> int:
        """rz_batch_need: bytes this context would hold after a batch of the
        given rz_batch_item records (nothing is recorded)."""
        items = np.ascontiguousarray(items, dtype=L.batch_item_dtype())
        out = C.c_uint64(0)
        self._check(self._L.rz_batch_need(self.handle, items.ctypes.data_as(C.c_void_p), len(items),
                                          C.byref(out)))
        return int(out.value)

    def trim(self):
        """rz_trim: free this context's pipeline buffers (between jobs)."""
        self._check(self._L.rz_trim(self.handle))

    def set_fine_tile(self, side: int):
        """Side (8, 16 or 32) of the tiles whose lists binning builds."""
        self._check(self._L.rz_set_fine_tile(self.handle, int(side)))

    def fine_tile(self) -> int:
        return int(self._L.rz_get_fine_tile(self.handle))

    def service_start(self, wg_per_cu: int = 2, idle_us: int = 50_000):
        """rz_service_start: launch the group's persistent raster service
        (wg_per_cu resident workgroups per CU; it exits after idle_us
        without work and is relaunched when work is posted)."""
        self._check(self._L.rz_service_start(self.handle, int(wg_per_cu), int(idle_us)))

    def service_stop(self):
        """rz_service_stop: stop the group's raster service (no attached
        context may have a job in flight)."""
        self._check(self._L.rz_service_stop(self.handle))

    def service_attach(self, on: bool = True):
        """rz_service_attach: this context's batches post their raster to the
        group's service (on) or run it themselves (off)."""
        self._check(self._L.rz_service_attach(self.handle, 1 if on else 0))

    def service_stats(self) -> dict:
        a = (C.c_uint64 * 4)()
        self._check(self._L.rz_service_stats(self.handle, a))
        return dict(on=bool(a[0]), running=bool(a[1]), launches=int(a[2]), grid=int(a[3]))

    def set_coarse_bin(self, pixels: int):
        self._check(self._L.rz_set_coarse_bin(self.handle, pixels))

    # ---- meshes
    def mesh(self, pos, color=None, uv=None, idx=None) -> DeviceMesh:
        pos = _f32(pos, 3)
        nv = pos.shape[0]
        color = None if color is None else _f32(color, 4)
        uv = None if uv is None else _f32(uv, 2)
        if idx is None:
            idx = np.arange(nv, dtype=np.uint32)
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint32).reshape(-1))
        h = C.c_void_p()
        self._check(self._L.rz_mesh_create(self.handle, _ptr(pos), _ptr(color), _ptr(uv), nv, _ptr(idx),
                                           idx.size, C.byref(h)))
        return DeviceMesh(self, h, idx.size // 3)

    def mesh_device(self, pos, color, uv, idx) -> DeviceMesh:
        """From torch device tensors (pos (nv,3) f32, color (nv,4) or None, uv (nv,2) or None, idx u32/int32)."""
        nv = int(pos.shape[0])
        dp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        h = C.c_void_p()
        self._check(self._L.rz_mesh_create_device(self.handle, dp(pos), dp(color), dp(uv), nv, dp(idx),
                                                  int(idx.numel()), C.byref(h)))
        return DeviceMesh(self, h, int(idx.numel()) // 3)

    # ---- frames
    def draw(self, mesh: DeviceMesh, fs: int, vs: int = L.RZ_VS_MVP):
        self._check(self._L.rz_draw(self.handle, mesh.handle, int(vs), int(fs)))

    def resolve(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.uint32)
        self._check(self._L.rz_resolve(self.handle, _ptr(out)))
        return out

    def resolve_device(self, dev_ptr: int):
        self._check(self._L.rz_resolve_device(self.handle, C.c_void_p(dev_ptr)))

    def resolve_msaa(self):
        out = np.empty((self.height, self.width), np.uint32)
        col = np.empty((self.height, self.width, 4), np.uint32)
        dep = np.empty((self.height, self.width, 4), np.float32)
        self._check(self._L.rz_resolve_msaa(self.handle, _ptr(out), _ptr(col), _ptr(dep)))
        return out, col, dep

    def frame_flags(self) -> int:
        return int(self._L.rz_frame_flags(self.handle))

    def frame_stats(self) -> dict:
        a = (C.c_uint64 * 11)()
        self._check(self._L.rz_frame_stats(self.handle, a))
        keys = ["tris", "clipped", "prims", "bin_entries", "replays", "raster_entries", "raster_candidates",
                "raster_prims", "raster_covered_samples", "raster_shading_tasks", "overflow_bits"]
        return dict(zip(keys, (int(x) for x in a)))

    # ---- batches
    def batch_begin(self):
        self._check(self._L.rz_batch_begin(self.handle))

    def batch_scene(self, width: int, height: int, clear_depth: float, dev_ptr: int) -> int:
        idx = C.c_uint32()
        self._check(self._L.rz_batch_scene(self.handle, int(width), int(height), float(clear_depth),
                                           C.c_void_p(dev_ptr), C.byref(idx)))
        return idx.value

    def batch_scene_msaa(self, scene: int, color4_ptr: int, depth4_ptr: int):
        self._check(self._L.rz_batch_scene_msaa(self.handle, int(scene), C.c_void_p(color4_ptr or None),
                                                C.c_void_p(depth4_ptr or None)))

    def batch_draw(self, scene: int, mesh: DeviceMesh, world, view, proj, texture: int, fs: int,
                   vs: int = L.RZ_VS_MVP):
        w, v, p = (_f32(m).reshape(16) for m in (world, view, proj))
        fp = C.POINTER(C.c_float)
        self._check(self._L.rz_batch_draw(self.handle, int(scene), mesh.handle, w.ctypes.data_as(fp),
                                          v.ctypes.data_as(fp), p.ctypes.data_as(fp), int(texture), int(vs),
                                          int(fs)))

    def batch_add(self, items) -> int:
        """items: numpy array of _lib.batch_item_dtype() records. Returns the
        index of the first scene added."""
        import numpy as np

        items = np.ascontiguousarray(items, dtype=L.batch_item_dtype())
        first = C.c_uint32()
        self._check(self._L.rz_batch_add(self.handle, C.c_void_p(items.ctypes.data), len(items), C.byref(first)))
        return first.value

    def batch_submit(self):
        self._check(self._L.rz_batch_submit(self.handle))

    def batch_query(self) -> bool:
        rc = self._L.rz_batch_query(self.handle)
        if rc < 0:
            self._check(rc)
        return rc == 1

    def batch_wait(self):
        self._check(self._L.rz_batch_wait(self.handle))

    def batch_scene_flags(self, scene: int) -> int:
        f = C.c_uint32()
        self._check(self._L.rz_batch_scene_flags(self.handle, int(scene), C.byref(f)))
        return f.value

    def batch_tuning(self, bin_target: int = 256, raster_per_cu: int = 0, use_graph: bool = True,
                     fine_tile: int | None = None):
        """rz_batch_tuning, and rz_set_fine_tile when fine_tile is given."""
        self._check(self._L.rz_batch_tuning(self.handle, int(bin_target), int(raster_per_cu),
                                            1 if use_graph else 0))
        if fine_tile is not None:
            self.set_fine_tile(fine_tile)

    def get_batch_tuning(self) -> dict:
        a = (C.c_uint32 * 3)()
        self._check(self._L.rz_batch_get_tuning(self.handle, a))
        return dict(bin_target=int(a[0]), raster_per_cu=int(a[1]), use_graph=bool(a[2]),
                    fine_tile=self.fine_tile())

    def profile(self, enable: bool):
        self._check(self._L.rz_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        ms = (C.c_float * L.RZ_NSTAGES)()
        n = C.c_uint32()
        self._check(self._L.rz_profile_read(self.handle, ms, C.byref(n)))
        return dict(geometry=ms[0], binning=ms[1], raster=ms[2], frames=n.value)


# ---------------------------------------------------------------- mirror API
class Texture:
    """src/texture.rs: RGBA8, (0, 0) is the upper-left texel."""

    def __init__(self, rgba8: np.ndarray):
        self.buf = np.ascontiguousarray(rgba8, dtype=np.uint8)
        if self.buf.ndim != 3 or self.buf.shape[2] != 4:
            raise ValueError("texture must be (height, width, 4) RGBA8")
        self.height, self.width = self.buf.shape[:2]

    @staticmethod
    def from_rgba8(rgba8) -> "Texture":
        return Texture(rgba8)

    @staticmethod
    def from_raw_file(path: str, width: int, height: int) -> "Texture":
        """A texture from a file of raw RGBA8 texels, row-major from the
        upper-left texel: the stand-in for the PNG decode of
        src/texture.rs:26-45 (the decoded pixel layout it produces)."""
        if width <= 0 or height <= 0:
            raise ValueError(f"texture size must be positive, got {width} x {height}")
        data = np.fromfile(path, dtype=np.uint8)
        if data.size != width * height * 4:
            raise ValueError(f"{path}: {data.size} bytes, expected {width} x {height} x 4")
        return Texture(data.reshape(height, width, 4))


class Mesh:
    """src/mesh.rs:5-12. vertices (n,3) f32, indices (3k,) int, attributes
    as colors (n,4) f32 and uvs (n,2) f32 (VertexAttribute, graphics_primitives.rs:9-13).
    Treat as immutable once rendered, or call invalidate()."""

    def __init__(self, vertices, indices, colors=None, uvs=None):
        self.vertices = _f32(vertices, 3)
        n = self.vertices.shape[0]
        self.indices = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
        self.colors = _f32(colors if colors is not None else np.zeros((n, 4)), 4)
        self.uvs = _f32(uvs if uvs is not None else np.zeros((n, 2)), 2)
        self.version = 0

    def invalidate(self):
        self.version += 1


class UniformBlock:
    """src/uniform.rs:5-9 (row-major 4x4 f32 matrices)."""

    def __init__(self):
        self.world = np.eye(4, dtype=np.float32)
        self.view = np.eye(4, dtype=np.float32)
        self.projection = np.eye(4, dtype=np.float32)


class Uniforms:
    """src/uniform.rs:12-46."""

    def __init__(self, ctx: Context):
        self._ctx = ctx
        self.textures: list[Texture] = []
        self.uniform_block = UniformBlock()

    def bind_texture(self, index: int, tex: Texture):
        # uniform.rs:29-33 asserts index == textures.len()
        if index != len(self.textures):
            raise ValueError("texture index must equal the number of bound textures")
        self._ctx.bind_texture(index, tex.buf)
        self.textures.append(tex)

    def get_texture(self, index: int) -> Texture:
        return self.textures[index]

    def read_block(self) -> UniformBlock:
        return self.uniform_block

    def write_block(self) -> UniformBlock:
        return self.uniform_block


class Renderer:
    """src/render.rs:38-145 without the window: `framebuffer()` returns the
    resolved ARGB image that `display()` would hand to minifb."""

    def __init__(self, width: int, height: int):
        self.width, self.height = width, height
        self._ctx = Context(width, height)
        self._uniforms = Uniforms(self._ctx)
        self._meshes: dict[int, tuple] = {}
        # device meshes replaced (Mesh.invalidate) or whose Mesh is gone; a
        # draw of the current frame may still use them, so they are
        # destroyed after the next framebuffer()
        self._retired: list[DeviceMesh] = []

    @property
    def context(self) -> Context:
        return self._ctx

    def uniforms(self) -> Uniforms:
        return self._uniforms

    def _device_mesh(self, mesh: Mesh) -> DeviceMesh:
        key = id(mesh)
        hit = self._meshes.pop(key, None)
        if hit is not None:
            ref, ver, dm = hit
            if ref() is mesh and ver == mesh.version:
                self._meshes[key] = hit
                return dm
            self._retired.append(dm)
        dm = self._ctx.mesh(mesh.vertices, mesh.colors, mesh.uvs, mesh.indices)
        meshes, retired = self._meshes, self._retired

        def gone(r, key=key):  # the Mesh was collected: retire its device copy
            e = meshes.get(key)
            if e is not None and e[0] is r:
                del meshes[key]
                retired.append(e[2])

        self._meshes[key] = (weakref.ref(mesh, gone), mesh.version, dm)
        return dm

    def _free_retired(self):
        while self._retired:
            self._retired.pop().destroy()

    def device_meshes(self) -> int:
        """Device meshes held (current and retired): for leak checks."""
        return len(self._meshes) + len(self._retired)

    def render(self, mesh: Mesh, vertex_shader: VertexShader = VertexShader.MVP,
               fragment_shader: FragmentShader = FragmentShader.TEXTURE):
        """render.rs:98-114: shade vertices with the current uniforms, assemble
        and rasterize; accumulates into the frame until framebuffer()."""
        b = self._uniforms.uniform_block
        self._ctx.set_uniforms(b.world, b.view, b.projection)
        self._ctx.draw(self._device_mesh(mesh), int(fragment_shader), int(vertex_shader))

    def framebuffer(self) -> np.ndarray:
        """Rasterizer::framebuffer (mod.rs:520-522): resolve + clear; (H, W) u32 ARGB."""
        img = self._ctx.resolve()
        self._free_retired()
        return img

    def close(self):
        for dm in [e[2] for e in self._meshes.values()] + self._retired:
            try:
                dm.destroy()
            except RzError:
                pass
        self._meshes.clear()
        self._retired.clear()
        self._ctx.close()
