"""Host-side f32 math that builds the per-draw uniforms.

The reference builds its matrices on the CPU once per frame
(src/math/{vector,matrix,transform,mod}.rs, src/camera.rs); the pipeline
only ever sees the resulting f32 bits. This module reproduces those bits:
every operation is rounded to f32 in the reference's order, dot products
accumulate from 0.0 left to right (vector.rs:17-23), and sin/cos/tan/sqrt
come from the C library's float functions, as Rust's f32 methods do.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

F = np.float32

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf", "tanf", "sqrtf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]


def sinf(x) -> np.float32:
    return F(_libm.sinf(float(F(x))))


def cosf(x) -> np.float32:
    return F(_libm.cosf(float(F(x))))


def tanf(x) -> np.float32:
    return F(_libm.tanf(float(F(x))))


def sqrtf(x) -> np.float32:
    return F(_libm.sqrtf(float(F(x))))


FRAC_PI_2 = F(np.pi / 2)
FRAC_PI_4 = F(np.pi / 4)
PI = F(np.pi)


# ------------------------------------------------------------ vectors
def dot(a, b) -> np.float32:
    """vector.rs:17-23."""
    s = F(0.0)
    for x, y in zip(a, b):
        s = F(s + F(F(x) * F(y)))
    return s


def length(v) -> np.float32:
    """vector.rs:38-40: fold(0.0, acc + e*e).sqrt()."""
    acc = F(0.0)
    for e in v:
        acc = F(acc + F(F(e) * F(e)))
    return sqrtf(acc)


def normalized(v) -> np.ndarray:
    """vector.rs:42-44 (component-wise divide, vector.rs:110-122)."""
    n = length(v)
    return np.array([F(F(e) / n) for e in v], dtype=np.float32)


def neg(v) -> np.ndarray:
    """vector.rs:82-94: each component *= -1.0."""
    return np.array([F(F(e) * F(-1.0)) for e in v], dtype=np.float32)


def cross3(a, b) -> np.ndarray:
    """vector.rs:196-204."""
    a = [F(x) for x in a]
    b = [F(x) for x in b]
    x = F(F(a[1] * b[2]) - F(a[2] * b[1]))
    y = F(F(a[2] * b[0]) - F(a[0] * b[2]))
    z = F(F(a[0] * b[1]) - F(a[1] * b[0]))
    return np.array([x, y, z], dtype=np.float32)


# ------------------------------------------------------------ matrices
def mat4(*vals) -> np.ndarray:
    """matrix.rs:16-41 (row-major)."""
    assert len(vals) == 16
    return np.array([F(v) for v in vals], dtype=np.float32).reshape(4, 4)


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


def transpose(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.float32).T)


def mat_mul(a, b) -> np.ndarray:
    """matrix.rs:56-79: result[i][j] = row_i(a) . col_j(b)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    out = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            out[i, j] = dot(a[i, :], b[:, j])
    return out


def mat_vec(m, v) -> np.ndarray:
    """vector.rs:219-240."""
    m = np.asarray(m, dtype=np.float32)
    return np.array([dot(m[i, :], v) for i in range(4)], dtype=np.float32)


def chain(*ms) -> np.ndarray:
    """Left-to-right product a * b * c ... as Rust evaluates it."""
    out = np.asarray(ms[0], dtype=np.float32)
    for m in ms[1:]:
        out = mat_mul(out, m)
    return out


# ------------------------------------------------------------ transforms
def translate(x, y, z) -> np.ndarray:
    """transform.rs:10-17."""
    return mat4(1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1)


def translate_v(v) -> np.ndarray:
    return translate(v[0], v[1], v[2])


def rotate_x(rad) -> np.ndarray:
    """transform.rs:19-41."""
    c, s = cosf(rad), sinf(rad)
    return mat4(1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1)


def rotate_y(rad) -> np.ndarray:
    """transform.rs:43-65."""
    c, s = cosf(rad), sinf(rad)
    return mat4(c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1)


def rotate_z(rad) -> np.ndarray:
    """transform.rs:67-89."""
    c, s = cosf(rad), sinf(rad)
    return mat4(c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)


def rotate(x, y, z) -> np.ndarray:
    """transform.rs:91-96: rotate_z(z) * rotate_y(y) * rotate_x(x)."""
    return chain(rotate_z(z), rotate_y(y), rotate_x(x))


def project(near, far, aspect_ratio, vert_fov) -> np.ndarray:
    """math/mod.rs:92-122 (vert_fov sets the horizontal half-width there)."""
    near, far, aspect_ratio, vert_fov = F(near), F(far), F(aspect_ratio), F(vert_fov)
    assert near > 0
    half_width = F(tanf(F(vert_fov / F(2.0))) * near)
    half_height = F(aspect_ratio * half_width)
    return mat4(
        F(near / half_width), 0, 0, 0,
        0, F(near / half_height), 0, 0,
        0, 0, F(F(-F(far + near)) / F(far - near)), F(F(F(F(-2.0) * far) * near) / F(far - near)),
        0, 0, -1, 0,
    )


class Camera:
    """src/camera.rs: position, up and look direction."""

    def __init__(self, pos=(0.0, 0.0, -5.0), up=(0.0, 1.0, 0.0), dir=(0.0, 0.0, 1.0)):
        self.pos = np.array(pos, dtype=np.float32)
        self.up = normalized(np.array(up, dtype=np.float32))
        self.dir = normalized(np.array(dir, dtype=np.float32))

    def get_view_matrix(self) -> np.ndarray:
        """camera.rs:10-43."""
        cam_z = neg(self.dir)
        cam_x = normalized(cross3(cam_z, self.up))
        cam_y = normalized(cross3(cam_x, cam_z))
        rotation_inv = transpose(mat4(
            cam_x[0], cam_y[0], cam_z[0], 0,
            cam_x[1], cam_y[1], cam_z[1], 0,
            cam_x[2], cam_y[2], cam_z[2], 0,
            0, 0, 0, 1))
        vec_to_pos = np.array([F(self.pos[i] - F(0.0)) for i in range(3)], dtype=np.float32)
        translation_inv = translate_v(neg(vec_to_pos))
        return mat_mul(rotation_inv, translation_inv)
