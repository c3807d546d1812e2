"""Seeded synthetic workloads (vertex streams, uniforms, textures).

  cube_scene     unit cube, 800x600, 256^2 checkerboard, camera inside the
                 near-clip range so several faces cross the near plane
  soup_scene     triangle soup: centroids uniform in the view frustum,
                 screen areas log-uniform in [0.5, 5000] px^2, front-facing,
                 a fixed share with 1-2 vertices behind the near plane
  batch_scenes   many soups of uneven size and resolution, each textured
  terrain_scene  subdivided fBm height field seen by a low camera

Arrays are f32 / u32 exactly as uploaded; matrices come from hostmath,
which reproduces the reference's f32 bits.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import hostmath as hm
from . import meshes

F = np.float32
FS_TEXTURE, FS_COLOR, FS_DEBUG = 0, 1, 2
CLEAR_DEPTH_MAX = float(np.finfo(np.float32).max)

SEEDS = {"soup": 1001, "batch": 1002, "terrain": 1003, "serve": 1004}


@dataclasses.dataclass
class Scene:
    name: str
    width: int
    height: int
    pos: np.ndarray            # (nv, 3) f32
    color: np.ndarray | None   # (nv, 4) f32
    uv: np.ndarray | None      # (nv, 2) f32
    idx: np.ndarray            # (ni,) u32
    world: np.ndarray
    view: np.ndarray
    proj: np.ndarray
    fs: int
    texture: np.ndarray | None = None  # (th, tw, 4) u8
    clear_depth: float = CLEAR_DEPTH_MAX

    @property
    def ntri(self) -> int:
        return int(self.idx.size // 3)

    @property
    def samples(self) -> int:
        return 4 * self.width * self.height

    def algorithmic_bytes(self) -> int:
        """Bytes the draw must move at least once: positions, the attributes
        the fragment shader reads, indices, the texture, the resolved image."""
        nv = self.pos.shape[0]
        attr = {FS_TEXTURE: 8, FS_COLOR: 16, FS_DEBUG: 0}[self.fs]
        tex = self.texture.size if (self.fs == FS_TEXTURE and self.texture is not None) else 0
        return nv * (12 + attr) + 4 * self.idx.size + tex + 4 * self.width * self.height


def _camera_at_origin():
    return hm.Camera(pos=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), dir=(0.0, 0.0, 1.0)).get_view_matrix()


# ------------------------------------------------------------------ cube
def cube_scene(width: int = 800, height: int = 600, clear_depth: float = CLEAR_DEPTH_MAX) -> Scene:
    """mesh::cube(1.0), textured, 60 degree projection; the camera sits close
    enough that the near plane cuts the cube. clear_depth: f32::MAX as the
    reference clears (buffers.rs:6), or 1.0 as the plain cube workload
    specifies."""
    m = meshes.cube(1.0)
    view = hm.Camera(pos=(0.15, 0.1, -0.8), up=(0.0, 1.0, 0.0), dir=(0.1, 0.05, 1.0)).get_view_matrix()
    proj = hm.project(0.45, 50.0, F(height) / F(width), F(math.pi / 3))
    world = hm.rotate(F(0.5), F(0.7), F(0.2))
    tex = meshes.checkerboard(256, 8)
    return Scene("cube", width, height, m.vertices, m.colors, m.uvs, m.indices, world, view, proj,
                 FS_TEXTURE, tex, clear_depth)


# ------------------------------------------------------------------ soup
def _soup_arrays(rng, n, width, height, near, far, tan_h, clip_frac, area_lo=0.5, area_hi=5000.0):
    aspect = height / width
    # centroid depth: uniform in the frustum volume (density ~ z^2)
    z_lo, z_hi = near * 1.05, far * 0.9
    u = rng.random(n)
    zc = np.cbrt(z_lo ** 3 + u * (z_hi ** 3 - z_lo ** 3))
    ux = rng.uniform(-1, 1, n)
    uy = rng.uniform(-1, 1, n)
    area = np.exp(rng.uniform(np.log(area_lo), np.log(area_hi), n))  # px^2
    theta0 = rng.uniform(0, 2 * np.pi, n)
    ang = theta0[:, None] + np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])[None, :] + rng.uniform(-0.4, 0.4, (n, 3))
    rad = rng.uniform(0.6, 1.4, (n, 3))
    sx, sy = rad * np.cos(ang), rad * np.sin(ang)
    shape_area = 0.5 * np.abs((sx[:, 1] - sx[:, 0]) * (sy[:, 2] - sy[:, 0]) - (sx[:, 2] - sx[:, 0]) * (sy[:, 1] - sy[:, 0]))
    crossing = np.zeros(n, bool)
    crossing[rng.permutation(n)[: int(round(clip_frac * n))]] = True
    # crossing triangles sit just beyond the near plane
    zc = np.where(crossing, near * rng.uniform(1.02, 1.6, n), zc)
    xc = ux * zc * tan_h
    yc = uy * zc * tan_h * aspect
    px_world = (2.0 / width) * zc * tan_h          # one pixel at depth zc
    k = np.sqrt(area / shape_area) * px_world       # world units per shape unit
    tilt = rng.uniform(-0.3, 0.3, (n, 3)) * k[:, None]
    P = np.empty((n, 3, 3))
    P[:, :, 0] = xc[:, None] + sx * k[:, None]
    P[:, :, 1] = yc[:, None] + sy * k[:, None]
    P[:, :, 2] = zc[:, None] + tilt
    # 1 or 2 vertices behind the near plane (w + z < 0 <=> z < near here)
    nback = rng.integers(1, 3, n)
    depth_back = near - rng.uniform(0.02, 1.0, (n, 3)) * np.maximum(k, 0.02 * near)[:, None]
    order = np.argsort(rng.random((n, 3)), axis=1)
    rank = np.argsort(order, axis=1)
    back = crossing[:, None] & (rank < nback[:, None])
    P[:, :, 2] = np.where(back, depth_back, P[:, :, 2])
    # front-facing: the reference covers samples only of triangles with a
    # positive screen-space 2x area (y down); seen from the eye at the origin
    # that is (p1 - p0) x (p2 - p0) . p0 < 0 here, also for the clipped part
    # of a triangle crossing the near plane (same plane, same side).
    nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    facing = np.einsum("ij,ij->i", nrm, P[:, 0])
    flip = facing > 0
    P[flip, 1], P[flip, 2] = P[flip, 2].copy(), P[flip, 1].copy()
    return P.astype(np.float32), flip


def soup_scene(n_tris: int = 1_000_000, width: int = 1920, height: int = 1080, seed: int = SEEDS["soup"],
               clip_frac: float = 0.10, textured: bool = False, near: float = 0.5, far: float = 50.0,
               texture: np.ndarray | None = None) -> Scene:
    rng = np.random.default_rng(seed)
    fov = math.pi / 2
    tan_h = float(hm.tanf(F(fov / 2)))
    P, flip = _soup_arrays(rng, n_tris, width, height, near, far, tan_h, clip_frac)
    pos = P.reshape(-1, 3)
    nv = pos.shape[0]
    color = rng.random((nv, 4), dtype=np.float32)
    uv = rng.random((nv, 2), dtype=np.float32) if textured else None
    idx = np.arange(nv, dtype=np.uint32)
    view = _camera_at_origin()
    proj = hm.project(F(near), F(far), F(height) / F(width), F(fov))
    world = hm.identity()
    fs = FS_TEXTURE if textured else FS_COLOR
    if textured and texture is None:
        texture = meshes.checkerboard(512, 8)
    return Scene(f"soup{n_tris}", width, height, pos, color, uv, idx, world, view, proj, fs,
                 texture if textured else None)


# ------------------------------------------------------------------ batch
RESOLUTIONS = (256, 512, 768, 1024)


def batch_scenes(n_scenes: int = 256, seed: int = SEEDS["batch"], tri_lo: int = 1000, tri_hi: int = 200_000,
                 resolutions=RESOLUTIONS, tex_size: int = 512):
    """Independent textured soups of uneven size and resolution, each with
    its own checkerboard palette."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        res = int(resolutions[rng.integers(0, len(resolutions))])
        nt = int(round(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi)))))
        pal = rng.integers(0, 256, (2, 4), dtype=np.uint8)
        pal[:, 3] = 255
        tex = meshes.checkerboard(tex_size, 8, tuple(pal[0]), tuple(pal[1]))
        sc = soup_scene(nt, res, res, seed=int(rng.integers(0, 2**31)), textured=True, texture=tex)
        sc.name = f"batch{s}"
        out.append(sc)
    return out


# ------------------------------------------------------------------ terrain
def _fbm(rng, n, octaves=6):
    """Value-noise fBm on an (n, n) grid, amplitude ~1."""
    h = np.zeros((n, n))
    amp = 1.0
    xs = np.linspace(0, 1, n)
    for o in range(octaves):
        cells = 4 * 2 ** o
        g = rng.standard_normal((cells + 1, cells + 1))
        t = xs * cells
        i0 = np.minimum(t.astype(int), cells - 1)
        f = t - i0
        f = f * f * (3 - 2 * f)
        a = g[i0][:, i0]
        b = g[i0][:, i0 + 1]
        c = g[i0 + 1][:, i0]
        d = g[i0 + 1][:, i0 + 1]
        row = a * (1 - f[None, :]) + b * f[None, :]
        row2 = c * (1 - f[None, :]) + d * f[None, :]
        h += amp * (row * (1 - f[:, None]) + row2 * f[:, None])
        amp *= 0.5
    return h


def terrain_scene(quads: int = 1415, width: int = 3840, height: int = 2160, seed: int = SEEDS["terrain"],
                  tex_size: int = 4096, detail: float = 0.038, height_scale: float = 0.1,
                  cam_height: float = 4.0) -> Scene:
    """A (quads x quads)-quad grid over [-1, 1] x [0, 2] (x, z) with seeded
    heights: 6-octave value-noise fBm plus a per-vertex detail octave of
    relative amplitude `detail`, all scaled by `height_scale`. uv spans the
    grid once; a tex_size^2 checkerboard (64 x 64 checks). The camera looks
    straight down from cam_height above the grid centre with a lens whose
    half-width at the ground is 1 unit (the grid fills the screen width), and
    its near plane lies at the terrain's mean level: the terrain above it is
    cut away, and the triangles the near plane slices through take the
    clipping path (about 10% of the full-size grid's 4,004,450 triangles)."""
    rng = np.random.default_rng(seed)
    n = quads + 1
    h = (_fbm(rng, n) * 0.08 + rng.uniform(-1.0, 1.0, (n, n)) * detail) * height_scale
    gx, gz = np.meshgrid(np.linspace(-1.0, 1.0, n), np.linspace(0.0, 2.0, n))
    pos = np.stack([gx, h, gz], axis=-1).reshape(-1, 3).astype(np.float32)
    uu, vv = np.meshgrid(np.linspace(0.0, 1.0, n), np.linspace(0.0, 1.0, n))
    uv = np.stack([uu, vv], axis=-1).reshape(-1, 2).astype(np.float32)
    i = np.arange(quads)
    a = (i[:, None] * n + i[None, :]).reshape(-1)  # row-major quad corners
    b, c, d = a + 1, a + n + 1, a + n
    # two triangles per quad, wound to face the camera above the terrain
    idx = np.stack([a, c, b, a, d, c], axis=-1).reshape(-1).astype(np.uint32)
    view = hm.Camera(pos=(0.0, cam_height, 1.0), up=(0.0, 0.0, 1.0), dir=(0.0, -1.0, 0.0)).get_view_matrix()
    fov = F(2.0 * math.atan(1.0 / cam_height))
    proj = hm.project(F(cam_height), F(2.0 * cam_height), F(height) / F(width), fov)
    tex = meshes.checkerboard(tex_size, 64)
    return Scene("terrain", width, height, pos, None, uv, idx, hm.identity(), view, proj, FS_TEXTURE, tex)


def nine_vertex_scene(n_tris: int = 200, width: int = 160, height: int = 96, seed: int = 1005) -> Scene:
    """Triangles that the frustum clips to 9-vertex polygons (clipping.rs:114-177
    with identity matrices, so clip space is the cube |x|, |y|, |z| <= w = 1).
    Each lies in a plane through the centre with normal near (1, 1, 1), which
    cuts the cube in a (near-)regular hexagon of circumradius sqrt(2); the
    triangle is equilateral about the centre with inradius r in (0.8, 1.3),
    its edge normals pointing at alternate hexagon corners, so each edge cuts
    one corner off (the neighbouring corners project to sqrt(2)/2 < r on the
    edge normal) and the 6 hexagon edges plus 3 triangle edges remain. The
    polygons outgrow the 8 lanes the clipper gives a polygon at first.
    Colour FS, random colours; the planes differ per triangle (rotation about
    the normal, inradius, a small tilt and offset), so the triangles overlap
    at distinct depths."""
    rng = np.random.default_rng(seed)
    tris = []
    for _ in range(n_tris):
        nrm = np.array([1.0, 1.0, 1.0]) + rng.uniform(-0.08, 0.08, 3)
        nrm /= np.linalg.norm(nrm)
        # corner directions of the hexagon, projected into the plane
        c0 = np.array([1.0, -1.0, 0.0])
        c0 -= nrm * (c0 @ nrm)
        c0 /= np.linalg.norm(c0)
        c1 = np.cross(nrm, c0)
        phi = rng.uniform(-0.15, 0.15)  # small turn: the edges still cut the same corners
        r = rng.uniform(0.8, 1.3)
        centre = nrm * rng.uniform(-0.05, 0.05)
        pts = []
        for k in range(3):
            a = phi + 2.0 * math.pi * k / 3.0  # edge normal k at angle a; its opposite vertex at a + pi
            d = -(math.cos(a) * c0 + math.sin(a) * c1)
            pts.append(centre + 2.0 * r * d)
        tris.append(pts)
    pos = np.array(tris, np.float64).reshape(-1, 3).astype(np.float32)
    n = pos.shape[0]
    color = rng.random((n, 4)).astype(np.float32)
    uv = np.zeros((n, 2), np.float32)
    idx = np.arange(n, dtype=np.uint32)
    return Scene(f"nine{n_tris}_{width}x{height}", width, height, pos, color, uv, idx,
                 hm.identity(), hm.identity(), hm.identity(), FS_COLOR)


def classify_fractions(sc: Scene) -> dict:
    """Shares of the scene's triangles by the reference's pre-clip
    classification (clipping.rs:62-112), from numpy f32 clip coordinates:
    degenerate, outside (trivial reject), inside (trivial accept), clipped."""
    M = np.asarray(hm.mat_mul(hm.mat_mul(sc.proj, sc.view), sc.world), np.float32).reshape(4, 4)
    p = np.concatenate([sc.pos, np.ones((len(sc.pos), 1), np.float32)], 1) @ M.T
    t = p[sc.idx.reshape(-1, 3)]
    x, y, z, w = t[..., 0], t[..., 1], t[..., 2], t[..., 3]
    a2 = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
    deg = np.abs(a2) < 1e-6
    inside = ((x >= -w) & (y >= -w) & (z >= -w) & (x <= w) & (y <= w) & (z <= w)).all(1)
    out = ((x < -w).all(1) | (y < -w).all(1) | (z < -w).all(1) | (x > w).all(1) | (y > w).all(1) | (z > w).all(1))
    return {"degenerate": float(deg.mean()), "outside": float((out & ~deg).mean()),
            "inside": float((inside & ~deg).mean()), "clipped": float((~deg & ~out & ~inside).mean())}
