"""Streamed terrain world on the CPU: terrain generator, pager and terrain scene.

The generator keeps src/gen/terrain.rs's structure (ball of radius 50, dirt band, stone,
tone = 0.5 * billow + 0.5, chunk layout x fastest, empty chunk -> None); its noise-rs
OpenSimplex/Billow samples are replaced by value noise (the crate is not available), so the
noise itself has no reference values: it is pinned at lattice points, where the interpolation
weights vanish and every sample reduces to the lattice hash (closed form below).
"""
import math

import numpy as np
import pytest

from vtrace_amd import renderer as R
from vtrace_amd import world as W

M32 = 0xFFFFFFFF


def pcg(v):
    v &= M32
    state = (v * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def lattice(x, y, z, s):
    h = pcg((x & M32) + pcg((y & M32) + pcg((z & M32) + pcg(s))))
    return 2.0 * ((h >> 8) * (1.0 / 16777216.0)) - 1.0


@pytest.fixture(scope="module")
def gen():
    return W.TerrainGenerator(0)


def test_ball_is_the_only_occupied_region(gen):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(-80, 81, 3))
        c = gen.voxel(x, y, z)
        if x * x + y * y + z * z > 2500:
            assert c == 0, (x, y, z)
        elif c:
            assert c >> 24 == 255


def test_signal_and_tone_at_lattice_points(gen):
    """voxel = 10 k: 0.1 * voxel is (within an ulp of) the integer k, so value() is the lattice
    value at k, and billow's octaves sample the lattice at 2^o k."""
    checked = 0
    for X in range(-5, 6):
        for Y in range(-5, 6):
            for Z in range(-5, 6):
                x, y, z = 10 * X, 10 * Y, 10 * Z
                if x * x + y * y + z * z > 2500:
                    continue
                c = gen.voxel(x, y, z)
                sig = lattice(X, Y, Z, 0)
                assert (c != 0) == (sig > 0), (x, y, z, sig)
                if not c:
                    continue
                s = 0.0
                norm = 0.0
                amp = 1.0
                for o in range(6):
                    v = lattice(X << o, Y << o, Z << o, 1 + o)
                    s += amp * (2.0 * abs(v) - 1.0)
                    norm += amp
                    amp *= 0.5
                tone = 0.5 * (s / norm) + 0.5
                surface = x * x + (y - 4) ** 2 + z * z > 2500
                base = (255.0, 200.0, 100.0) if surface else (150.0, 150.0, 150.0)
                want = [min(255, int(math.floor(tone * b))) for b in base]
                got = [(c >> (8 * k)) & 255 for k in range(3)]
                # the samples sit within an ulp of the lattice point: allow one step of rounding
                assert all(abs(g - w) <= 1 for g, w in zip(got, want)), (x, y, z, got, want)
                checked += 1
    assert checked > 100


def test_colour_bands(gen):
    """Dirt (255:200:100) where (x, y-4, z) leaves the ball, grey stone elsewhere."""
    dirt = stone = 0
    for cy in (-4, -3, 2, 3):
        rgba = np.zeros((16, 16, 16), np.uint32)
        if gen.lib.vt_terrain_chunk(0, cy, 0, 0, rgba.ctypes.data) != 1:
            continue
        zz, yy, xx = np.nonzero(rgba)
        for z, y, x in zip(zz, yy, xx):
            wx, wy, wz = int(x), int(y) + 16 * cy, int(z)
            c = int(rgba[z, y, x])
            r, g, b = c & 255, (c >> 8) & 255, (c >> 16) & 255
            if wx * wx + (wy - 4) ** 2 + wz * wz > 2500:
                # r = floor(255 t), g = floor(200 t), b = floor(100 t) for one t in [0, 1]
                lo = max(r / 255.0, g / 200.0, b / 100.0)
                hi = min((r + 1) / 255.0, (g + 1) / 200.0, (b + 1) / 100.0)
                assert lo < hi + 1e-12, (wx, wy, wz, r, g, b)
                dirt += 1
            else:
                assert r == g == b
                stone += 1
    assert dirt > 0 and stone > 0


def test_chunk_layout_x_fastest(gen):
    """terrain.rs:71-84: chunk voxel (x, y, z) = world voxel 16c + (x, y, z), stored x fastest."""
    for c in ((0, 0, 0), (-1, 2, -3), (2, -1, 1)):
        v = gen.gen_chunk(*c)
        assert v is not None
        d = v.data
        assert d.shape == (16, 16, 16) and d.dtype == np.uint32
        for (z, y, x) in ((0, 0, 0), (1, 2, 3), (15, 0, 7), (4, 15, 15), (9, 9, 1)):
            assert int(d[z, y, x]) == gen.voxel(16 * c[0] + x, 16 * c[1] + y, 16 * c[2] + z)


def test_empty_chunks_and_batch(gen):
    assert gen.gen_chunk(10, 0, 0) is None          # outside the ball
    assert gen.gen_chunk(0, 0, 0) is not None
    pos = [(x, y, z) for x in (-4, -1, 0, 3) for y in (-4, 0, 3) for z in (-2, 1)]
    batch = gen.gen_chunks(pos)
    for p, b in zip(pos, batch):
        one = gen.gen_chunk(*p)
        assert (one is None) == (b is None)
        if one is not None:
            assert np.array_equal(one.data, b.data)


def test_get_chunk_pos_floors():
    assert W.get_chunk_pos((0.0, 0.0, 0.0)) == (0, 0, 0)
    assert W.get_chunk_pos((1.99, -0.01, -2.0)) == (0, -1, -1)
    assert W.get_chunk_pos((-2.01, 4.0, 3.9)) == (-2, 2, 1)


def test_pager_caches_and_queues_once():
    q = R.TextureUploadQueue()
    p = W.WorldPager(0)
    h = p.page(0, 0, 0, q)
    assert h == 0 and len(q.queue) == 1
    assert p.page(0, 0, 0, q) == h and len(q.queue) == 1     # pager.rs:60
    assert p.page(9, 9, 9, q) is None and (9, 9, 9) in p.chunks  # pager.rs:72-74
    assert p.page(9, 9, 9, q) is None and len(q.queue) == 1


def test_page_many_equals_sequential_pages():
    pos = W.load_positions((0.3, -1.0, 2.5))
    q1, q2 = R.TextureUploadQueue(), R.TextureUploadQueue()
    a = W.WorldPager(0)
    b = W.WorldPager(0)
    h1 = [a.page(*p, q1) for p in pos[:3000]]
    h2 = b.page_many(pos[:3000], q2)
    assert h1 == h2
    assert [h for _, h in q1.queue] == [h for _, h in q2.queue]
    for (v1, _), (v2, _) in zip(q1.queue, q2.queue):
        assert np.array_equal(v1.data, v2.data)


def test_terrain_scene_matches_world_loop():
    """world.rs:163-198: x outermost, translate(chunk * 2) * scale(2), one child per chunk."""
    cam = (1.0, 3.0, -5.0)
    q = R.TextureUploadQueue()
    pager = W.WorldPager(0)
    scene = W.terrain_scene(cam, pager, q)
    pos = W.load_positions(cam)
    assert len(pos) == 21 ** 3
    assert pos[0] == (0 - 10, 1 - 10, -3 - 10) and pos[1] == (-10, -9, -12)
    filled = [p for p in pos if pager.chunks[p] is not None]
    kids = list(scene)
    assert len(kids) == len(filled) == scene.num_total_children() == len(q.queue)
    for (model, handle), p in zip(kids, filled):
        M = model.reshape(4, 4)  # column-major: M[c][r]
        assert np.array_equal(np.diag(M)[:3], [2.0, 2.0, 2.0]) and M[3][3] == 1.0
        assert np.array_equal(M[3][:3], np.float32(p) * np.float32(2.0))
        assert handle == pager.chunks[p][1]
    # paging the same view again generates and queues nothing new
    n = len(q.queue)
    again = W.terrain_scene(cam, pager, q)
    assert len(q.queue) == n and [h for _, h in again] == [h for _, h in scene]
