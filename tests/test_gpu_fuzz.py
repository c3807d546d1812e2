"""Randomized frames on a 1/8-pixel lattice, GPU against the oracle bit for
bit. With identity matrices and power-of-two screens every vertex lands
exactly on the lattice the sample offsets live on (multiples of 1/8 px), so
edges run through sample points (the tie rule, mod.rs:148-170), coverage
and depth sit on the exact boundaries of the raster's conservative culls,
and depths drawn from a few values tie exactly (strict LESS, mod.rs:363-378).
Vertices past the screen edges and past the near and far planes exercise
clipping (clipping.rs:114-194); repeated triangles, slivers, back-facing and
zero-area triangles, uv outside [0, 1] (texel bounds, texture.rs:47-63) and
the three clear depths and fine tile sides are mixed in."""
import numpy as np
import pytest

from rusterizer_amd import hostmath as hm
from rusterizer_amd import meshes, scenes
from tests.harness import assert_parity, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

F = np.float32
SIZES = [(16, 16), (32, 64), (64, 32), (64, 64), (128, 32), (128, 128)]
DEPTHS = F([0.25, 0.5, 0.5, 0.75, 1.0])
CLEARS = [scenes.CLEAR_DEPTH_MAX, 1.0, 0.5]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    from tests.conftest import gpu_available

    if not gpu_available():
        pytest.skip("no GPU")


def _lattice(rng, lo, hi, n):
    """n values on the 1/8 lattice in [lo, hi]."""
    return rng.integers(int(lo * 8), int(hi * 8) + 1, n).astype(np.float64) / 8.0


def snapped_frame(seed: int) -> scenes.Scene:
    rng = np.random.default_rng(seed)
    W, H = SIZES[seed % len(SIZES)]
    fs = (scenes.FS_COLOR, scenes.FS_TEXTURE, scenes.FS_DEBUG)[seed % 3]
    tris = []
    for _ in range(int(rng.integers(40, 160))):
        kind = rng.random()
        if kind < 0.45:  # small: within a few pixels
            cx, cy = _lattice(rng, -2, W + 2, 1)[0], _lattice(rng, -2, H + 2, 1)[0]
            p = [(cx + _lattice(rng, -3, 3, 1)[0], cy + _lattice(rng, -3, 3, 1)[0]) for _ in range(3)]
        elif kind < 0.8:  # anywhere, some far past the screen (clipped by the side planes)
            p = [(_lattice(rng, -W / 2, 1.5 * W, 1)[0], _lattice(rng, -H / 2, 1.5 * H, 1)[0]) for _ in range(3)]
        elif kind < 0.9:  # sliver: nearly or exactly collinear
            x0, y0 = _lattice(rng, 0, W, 1)[0], _lattice(rng, 0, H, 1)[0]
            dx, dy = _lattice(rng, -W / 2, W / 2, 1)[0], _lattice(rng, -H / 2, H / 2, 1)[0]
            k = int(rng.integers(0, 3))
            p = [(x0, y0), (x0 + dx, y0 + dy), (x0 + 2 * dx + k / 8.0, y0 + 2 * dy)]
        else:  # axis-aligned right triangle on pixel or sample lines
            x0, y0 = _lattice(rng, 0, W, 1)[0], _lattice(rng, 0, H, 1)[0]
            s = _lattice(rng, 1, 12, 1)[0]
            p = [(x0, y0), (x0 + s, y0), (x0, y0 + s)]
        z = rng.choice(DEPTHS, 3) if rng.random() < 0.7 else rng.random(3).astype(F)
        if rng.random() < 0.08:
            z[int(rng.integers(0, 3))] = F(-1.5)  # behind the near plane
        if rng.random() < 0.04:
            z[int(rng.integers(0, 3))] = F(1.5)  # past the far plane
        tri = [(sx / (W / 2) - 1.0, 1.0 - sy / (H / 2), float(zz)) for (sx, sy), zz in zip(p, z)]
        tris.append(tri)
        if rng.random() < 0.1:
            tris.append(tri)  # drawn again later: exact depth ties
        if rng.random() < 0.1:
            tris.append([tri[0], tri[2], tri[1]])  # the other winding
    pos = np.array(tris, F).reshape(-1, 3)
    n = pos.shape[0]
    col = rng.random((n, 4)).astype(F)
    uv = rng.random((n, 2)).astype(F)
    if fs == scenes.FS_TEXTURE:
        out = rng.random(n) < 0.05
        uv[out] = (rng.random((int(out.sum()), 2)) * 3.0 - 1.0).astype(F)  # past the texture
    tex = meshes.checkerboard(32, 4) if fs == scenes.FS_TEXTURE else None
    sc = scenes.Scene(f"snapped{seed}_{W}x{H}_fs{fs}", W, H, pos, col, uv, np.arange(n, dtype=np.uint32),
                      hm.identity(), hm.identity(), hm.identity(), fs, tex)
    sc.clear_depth = CLEARS[(seed // 3) % len(CLEARS)]
    return sc


def test_snapped_random_frames():
    drawn = 0
    sides = (8, 16, 32)
    for seed in range(180):
        sc = snapped_frame(seed)
        side = sides[(seed // 2) % 3]
        print(f"frame {sc.name} clear {sc.clear_depth} fine tile {side}", flush=True)
        g = run_gpu(sc, fine_tile=side)
        c = run_oracle(sc)
        assert_parity(g, c, where=f"{sc.name} fine tile {side}")
        drawn += int((c["col"] != 0xFF191919).sum()) > 0
    assert drawn >= 150  # nearly every frame draws something


@pytest.mark.parametrize("tuning", [dict(bin_target=16, raster_per_cu=0, use_graph=True, fine_tile=32),
                                    dict(bin_target=64, raster_per_cu=0, use_graph=False, fine_tile=16),
                                    dict(bin_target=256, raster_per_cu=8, use_graph=True, fine_tile=8)],
                         ids=["ft32_graph", "ft16_eager", "ft8_persistent"])
def test_snapped_frames_as_batch(tuning):
    """The same kind of frames, 60 of them as the scenes of one batch (each
    scene its own size, clear depth and shader), every scene bit-exact."""
    from tests.harness import run_batch

    scs = [snapped_frame(1000 + k) for k in range(60)]
    print(f"batch of {len(scs)} snapped scenes, tuning {tuning}", flush=True)
    res, _ = run_batch(scs, tuning=tuning)
    for sc, r in zip(scs, res):
        assert_parity(r, run_oracle(sc), where=f"batch {sc.name}")
