"""The reference's mesh builders (src/mesh.rs) and its demo texture.

Values are computed in f32 in the reference's order; sin/cos come from the
C library (hostmath), as Rust's f32 methods do.
"""
from __future__ import annotations

import numpy as np

from . import hostmath as hm
from .render import Mesh

F = np.float32
RED, GREEN, BLUE, WHITE = (1, 0, 0, 1), (0, 1, 0, 1), (0, 0, 1, 1), (1, 1, 1, 1)  # color.rs:32-63


def centered_quad(width) -> Mesh:
    """mesh.rs:15-40."""
    h = F(F(width) / F(2.0))
    v = [(-h, h, 2.0), (h, h, 2.0), (h, -h, 2.0), (-h, -h, 2.0)]
    cols = [RED, BLUE, GREEN, WHITE]
    uvs = [(0, 0), (1, 0), (1, 1), (0, 1)]
    return Mesh(v, [0, 1, 2, 0, 2, 3], cols, uvs)


def triangle() -> Mesh:
    """mesh.rs:43-65."""
    v = [(-1.0, -1.0, 2.0), (0.0, 1.0, 2.0), (1.0, -1.0, 2.0)]
    return Mesh(v, [0, 1, 2], [RED, BLUE, GREEN], [(0, 1), (1, 0), (1, 1)])


_CUBE = [
    (-0.5, 0.5, -0.5), (0.5, 0.5, -0.5), (0.5, -0.5, -0.5), (-0.5, -0.5, -0.5),  # front
    (0.5, 0.5, 0.5), (-0.5, 0.5, 0.5), (-0.5, -0.5, 0.5), (0.5, -0.5, 0.5),      # back
    (-0.5, 0.5, 0.5), (-0.5, 0.5, -0.5), (-0.5, -0.5, -0.5), (-0.5, -0.5, 0.5),  # left
    (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (0.5, -0.5, 0.5), (0.5, -0.5, -0.5),      # right
    (-0.5, 0.5, -0.5), (-0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0.5, 0.5, -0.5),      # top
    (-0.5, -0.5, 0.5), (-0.5, -0.5, -0.5), (0.5, -0.5, -0.5), (0.5, -0.5, 0.5),  # bottom
]


def cube(width) -> Mesh:
    """mesh.rs:68-149."""
    w = F(width)
    verts = np.array([[F(F(c) * w) for c in p] for p in _CUBE], dtype=np.float32)  # point.rs:99-106
    idx = []
    for i in range(6):
        o = i * 4
        idx += [o + 0, o + 1, o + 2, o + 0, o + 2, o + 3]
    cols = [(RED, BLUE, GREEN)[i % 3] for i in range(len(verts))]
    uvs = [(0, 0), (1, 0), (1, 1), (0, 1)] * (len(verts) // 4)
    return Mesh(verts, idx, cols, uvs)


def sphere(radius) -> Mesh:
    """mesh.rs:152-207 (17 x 9 samples, ISO spherical coordinates)."""
    n_phi, n_theta = 17, 9
    r = F(radius)
    verts, idx, cols, uvs = [], [], [], []
    for i in range(n_theta):
        for j in range(n_phi):
            theta_ratio = F(F(i) / F(n_theta - 1))
            phi_ratio = F(F(j) / F(n_phi - 1))
            phi = F(F(hm.PI * F(2.0)) * phi_ratio)
            theta = F(hm.PI * theta_ratio)
            x = F(F(r * hm.sinf(theta)) * hm.cosf(phi))
            y = F(r * hm.cosf(theta))
            z = F(F(r * hm.sinf(theta)) * hm.sinf(phi))
            verts.append((x, y, z))
            if i < n_theta - 1 and j < n_phi - 1:
                idx += [n_phi * i + j, n_phi * i + (j + 1), n_phi * (i + 1) + (j + 1)]
                idx += [n_phi * i + j, n_phi * (i + 1) + (j + 1), n_phi * (i + 1) + j]
            cols.append((abs(x), abs(y), abs(z), 1.0))
            uvs.append((phi_ratio, theta_ratio))
    return Mesh(verts, idx, cols, uvs)


def checkerboard(size: int, checks: int, dark=(0, 0, 0, 255), light=(255, 255, 255, 255)) -> np.ndarray:
    """size x size RGBA8 board of checks x checks squares, (0, 0) dark."""
    cell = max(size // checks, 1)
    i = np.arange(size) // cell
    odd = (i[:, None] + i[None, :]) % 2 == 1
    out = np.empty((size, size, 4), np.uint8)
    out[~odd] = dark
    out[odd] = light
    return out


def demo_checkerboard() -> np.ndarray:
    """images/checkerboard.png of the reference: 400x400 RGBA8 with 100-texel
    checks, texel (x, y) black when x/100 + y/100 is even, else white."""
    return checkerboard(400, 4)
