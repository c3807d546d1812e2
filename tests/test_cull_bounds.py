"""Soundness of the raster's exact depth culls (CPU property test).

The raster skips work when a lower bound on the depth a triangle can
interpolate is not below the stored depth. Two bounds exist: zlo (from the
vertex depths) and a plane pz0 + pzx*(x-bx0) + pzy*(y-by0) whose f64
construction is restated here from k_setup. For random triangles -- ordinary,
slivers, huge, tiny, steep depth slopes -- and every covered sample in their
bbox, the reference's f32 depth (mod.rs:225-253, evaluated with numpy f32 in
the reference's order) must be >= both bounds, and so must the raster's f32
evaluation of the plane at the quadrant corner.
"""
import numpy as np
import pytest

F = np.float32
RGSS = np.array([[5 / 8, 1 / 8], [7 / 8, 5 / 8], [3 / 8, 7 / 8], [1 / 8, 3 / 8]], np.float32)


def clamp01(x):
    x = np.where(x < 0, F(0), x)
    return np.where(x > 1, F(1), x).astype(np.float32)


def setup(p, z):
    """Edge setup as mod.rs:187-222 (f32)."""
    n = []
    for k in range(3):
        q = p[(k + 1) % 3]
        vx = F(q[0] - p[k][0])
        vy = F(q[1] - p[k][1])
        n.append((F(-vy), vx))
    v10 = (F(p[1][0] - p[0][0]), F(p[1][1] - p[0][1]))
    v20 = (F(p[2][0] - p[0][0]), F(p[2][1] - p[0][1]))
    inv = F(F(1.0) / F(F(v10[0] * v20[1]) - F(v20[0] * v10[1])))
    return n, inv


def plane_bound(p, n, inv, z, bbox):
    """Restatement of depth_plane_bound() in k_setup (f64)."""
    u = 2.0 ** -24
    x0, x1, y0, y1 = bbox
    z0, z1, z2 = (float(v) for v in z)
    dz0, dz1 = z0 - z2, z1 - z2
    (n1x, n1y), (n2x, n2y) = (float(v) for v in n[1]), (float(v) for v in n[2])
    a1x, a1y = float(p[1][0]), float(p[1][1])
    a2x, a2y = float(p[2][0]), float(p[2][1])
    inv = float(inv)
    bx, by, bw, bh = float(x0), float(y0), float(x1 - x0), float(y1 - y0)
    px = inv * (n1x * dz0 + n2x * dz1)
    py = inv * (n1y * dz0 + n2y * dz1)
    e1 = n1x * (bx - a1x) + n1y * (by - a1y)
    e2 = n2x * (bx - a2x) + n2y * (by - a2y)
    zb = z2 + inv * e1 * dz0 + inv * e2 * dz1
    m1 = abs(n1x) * max(abs(bx - a1x), abs(bx + bw - a1x)) + abs(n1y) * max(abs(by - a1y), abs(by + bh - a1y))
    m2 = abs(n2x) * max(abs(bx - a2x), abs(bx + bw - a2x)) + abs(n2y) * max(abs(by - a2y), abs(by + bh - a2y))
    d = 10.0 * u * max(m1, m2) * abs(inv) + 2.0 * u
    slack = (3.0 * d + 16.0 * u) * (abs(z0) + abs(z1) + abs(z2))
    ev = 8.0 * u * (abs(zb) + abs(px) * bw + abs(py) * bh) + 1e-30
    lo = zb - 2.0 * (slack + ev)
    ok = x0 < x1 and y0 < y1 and np.isfinite([lo, px, py, zb]).all() and abs(px) < 1e30 and abs(py) < 1e30
    if not ok:
        return F(-3.4028235e38), F(0), F(0)
    return F(lo), F(px), F(py)


def zlo_of(z):
    zm = F(min(z))
    if zm >= 0:
        return F(F(zm * F(1 - 2 ** -18)) - F(2 ** -126))
    return F(F(zm * F(2 + 2 ** -17)) - F(2 ** -126))


def covered_depths(p, n, inv, z, bbox, tie):
    """All covered samples in the bbox: (x, y, depth) with the reference's f32 ops."""
    x0, x1, y0, y1 = bbox
    xs, ys = np.meshgrid(np.arange(x0, x1, dtype=np.float32), np.arange(y0, y1, dtype=np.float32))
    xs, ys = xs.ravel(), ys.ravel()
    out_x, out_y, out_z = [], [], []
    for s in range(4):
        sx = (xs + RGSS[s, 0]).astype(np.float32)
        sy = (ys + RGSS[s, 1]).astype(np.float32)
        e = []
        ins = np.ones(sx.shape, bool)
        for k in range(3):
            t1 = (F(n[k][0]) * (sx - F(p[k][0])).astype(np.float32)).astype(np.float32)
            t2 = (F(n[k][1]) * (sy - F(p[k][1])).astype(np.float32)).astype(np.float32)
            ek = ((F(0) + t1).astype(np.float32) + t2).astype(np.float32)
            e.append(ek)
            ins &= (ek > 0) | (~(ek < 0) & tie[k])
        b0 = clamp01((e[1] * inv).astype(np.float32))
        b1 = clamp01((e[2] * inv).astype(np.float32))
        b2 = clamp01(((F(1) - b0).astype(np.float32) - b1).astype(np.float32))
        zz = (((b0 * F(z[0])).astype(np.float32) + (b1 * F(z[1])).astype(np.float32)).astype(np.float32)
              + (b2 * F(z[2])).astype(np.float32)).astype(np.float32)
        out_x.append(sx[ins])
        out_y.append(sy[ins])
        out_z.append(zz[ins])
    return np.concatenate(out_x), np.concatenate(out_y), np.concatenate(out_z)


def random_triangle(rng, kind):
    c = rng.uniform(20, 200, 2)
    if kind == "normal":
        p = c + rng.uniform(-30, 30, (3, 2))
    elif kind == "sliver":
        d = rng.normal(size=2)
        d /= np.linalg.norm(d)
        t = rng.uniform(-60, 60, 3)
        off = rng.uniform(-0.05, 0.05, 3)
        p = c + t[:, None] * d + off[:, None] * np.array([-d[1], d[0]])
    elif kind == "tiny":
        p = c + rng.uniform(-0.8, 0.8, (3, 2))
    else:  # big
        p = c + rng.uniform(-150, 150, (3, 2))
    p = p.astype(np.float32)
    # clockwise (positive 2x area in the reference's convention)
    a = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
    if a < 0:
        p[[1, 2]] = p[[2, 1]]
    zc = rng.uniform(0.0, 1.0)
    spread = {"normal": 0.05, "sliver": 0.2, "tiny": 0.01, "big": 0.5}[kind] * rng.uniform(0, 1)
    z = np.clip(zc + rng.uniform(-spread, spread, 3), 0.0, 1.0).astype(np.float32)
    return p, z


@pytest.mark.parametrize("kind", ["normal", "sliver", "tiny", "big"])
def test_depth_bounds_hold(kind):
    rng = np.random.default_rng({"normal": 1, "sliver": 2, "tiny": 3, "big": 4}[kind])
    checked = 0
    for _ in range({"big": 60, "sliver": 800, "tiny": 3000}.get(kind, 200)):
        p, z = random_triangle(rng, kind)
        n, inv = setup(p, z)
        if not np.isfinite(inv):
            continue
        tie = [bool((nx > 0) or (not (nx < 0) and ny < 0)) for nx, ny in n]
        x0 = int(max(0, np.floor(p[:, 0].min())))
        x1 = int(np.ceil(p[:, 0].max()))
        y0 = int(max(0, np.floor(p[:, 1].min())))
        y1 = int(np.ceil(p[:, 1].max()))
        if x0 >= x1 or y0 >= y1:
            continue
        bbox = (x0, x1, y0, y1)
        pz0, pzx, pzy = plane_bound(p, n, inv, z, bbox)
        zlo = zlo_of(z)
        xs, ys, zz = covered_depths(p, n, inv, z, bbox, tie)
        if zz.size == 0:
            continue
        ok = np.isfinite(zz)
        xs, ys, zz = xs[ok], ys[ok], zz[ok]
        assert (zz >= zlo).all(), (kind, p, z)
        # the raster evaluates the plane in f32 at a sample position
        bound = (F(pz0) + (F(pzx) * (xs - F(x0))).astype(np.float32)).astype(np.float32)
        bound = (bound + (F(pzy) * (ys - F(y0))).astype(np.float32)).astype(np.float32)
        bad = zz < bound
        assert not bad.any(), (kind, p.tolist(), z.tolist(), float((bound - zz)[bad].max()))
        checked += zz.size
    assert checked > 500
