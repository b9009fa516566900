"""CPU: pins of the oracle against closed forms, brute force and invariants (no GPU).

The reference ships no golden vectors or tests for this path (its only tests are the two
container tests in src/voxel/rawchunk.rs:324-347), so every pin here is independent of the
oracle's own code: textbook definitions (bit interleave, matrix inverse), brute-force ray /
voxel intersection in float64, closed-form radiance, and structural invariants.
"""
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
from vtrace_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- hash RNG ------------------------------------------------------------------------------
def test_pcg_hash_is_a_bijection_on_a_large_subset():
    L = O.lib()
    xs = np.arange(0, 1 << 16, dtype=np.uint64) * np.uint64(65537)
    hs = np.array([L.orc_pcg_hash(int(x) & 0xFFFFFFFF) for x in xs[:20000]], dtype=np.uint64)
    assert len(np.unique(hs)) == len(hs)


def test_pcg_hash_committed_vectors_and_uniformity():
    L = O.lib()
    vec = json.load(open(os.path.join(GOLDEN, "rng_vectors.json")))
    for v, h in vec["pcg_hash"]:
        assert L.orc_pcg_hash(v) == h
    for (seed, frame, pixel, sample), h in vec["rng_seed"]:
        assert L.orc_rng_seed(seed, frame, pixel, sample) == h
    u = np.array([(L.orc_pcg_hash(i) >> 8) / 16777216.0 for i in range(20000)])
    assert abs(u.mean() - 0.5) < 0.01
    counts, _ = np.histogram(u, bins=16, range=(0, 1))
    chi2 = ((counts - 1250.0) ** 2 / 1250.0).sum()
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


def test_product_and_oracle_hash_agree():
    L = O.lib()
    for v in (0, 1, 2, 12345, 0xFFFFFFFF, 0x9E3779B9):
        assert L.orc_pcg_hash(v) == scenes.pcg_hash(v)


# ---- Morton --------------------------------------------------------------------------------
def interleave(x, y, z):
    m = 0
    for b in range(10):
        m |= ((x >> b) & 1) << (3 * b + 2)
        m |= ((y >> b) & 1) << (3 * b + 1)
        m |= ((z >> b) & 1) << (3 * b)
    return m


def test_morton_known_answers():
    L = O.lib()
    assert L.orc_morton30(1, 0, 0) == 0b100
    assert L.orc_morton30(0, 1, 0) == 0b010
    assert L.orc_morton30(0, 0, 1) == 0b001
    assert L.orc_morton30(1023, 1023, 1023) == (1 << 30) - 1
    assert L.orc_morton30(2, 0, 0) == 1 << 5
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, size=(300, 3)):
        assert L.orc_morton30(int(x), int(y), int(z)) == interleave(int(x), int(y), int(z))


# ---- inverse ---------------------------------------------------------------------------------
def test_inverse4_against_numpy():
    L = O.lib()
    rng = np.random.default_rng(2)
    for _ in range(200):
        A = np.eye(4)
        A[:3, :3] = rng.normal(size=(3, 3)) * rng.uniform(0.1, 100.0)
        A[:3, 3] = rng.normal(size=3) * 1000
        m = A.T.astype(np.float32).reshape(16)  # column-major
        out = np.zeros(16, dtype=np.float32)
        assert L.orc_inverse4(m.ctypes.data, out.ctypes.data) == 1
        ref = np.linalg.inv(m.reshape(4, 4).T.astype(np.float64))
        got = out.reshape(4, 4).T.astype(np.float64)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-6 * scale
    sing = np.zeros(16, dtype=np.float32)
    out = np.zeros(16, dtype=np.float32)
    assert L.orc_inverse4(sing.ctypes.data, out.ctypes.data) == 0


# ---- instances and LBVH ----------------------------------------------------------------------
def _random_records(n, seed, nvol=3):
    rng = np.random.default_rng(seed)
    recs = []
    for i in range(n):
        t = rng.uniform(-5000, 5000, size=3)
        s = rng.uniform(8, 200, size=3)
        recs.append(scenes.instance_record(scenes.model_matrix(tuple(t), tuple(s)), int(rng.integers(0, nvol + 1))))
    return np.stack(recs)


def test_instance_prep_aabb_contains_corners_and_texture_bits():
    recs = _random_records(64, 3)
    inv, aabb, tex, valid = O.instance_prep(recs, 3)
    for i, r in enumerate(recs):
        assert tex[i] == r.view(np.uint32)[15]
        assert valid[i] == (1 if tex[i] < 3 else 0)
        M = r.copy()
        M[15] = 1.0
        M = M.reshape(4, 4).T.astype(np.float64)
        for c in range(8):
            p = np.array([0.5 if c & 1 else -0.5, 0.5 if c & 2 else -0.5, 0.5 if c & 4 else -0.5, 1.0])
            w = M @ p
            assert np.all(w[:3] >= aabb[i, :3]) and np.all(w[:3] <= aabb[i, 3:])


@pytest.mark.parametrize("n", [1, 2, 7, 64, 500])
def test_lbvh_structure(n):
    recs = _random_records(n, 10 + n)
    if n >= 7:
        recs[4] = recs[2]  # duplicate Morton code
    nodes, order, codes = O.lbvh(recs, 3)
    inv, aabb, tex, valid = O.instance_prep(recs, 3)
    # Morton codes of the padded-AABB centroids, quantized to 1024 cells of the centroid bounds
    cen = ((aabb[:, :3] + aabb[:, 3:]) * np.float32(0.5)).astype(np.float32)
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    ext = (hi - lo).astype(np.float32)
    q = np.zeros_like(cen, dtype=np.int64)
    for a in range(3):
        if ext[a] > 0:
            u = ((cen[:, a] - lo[a]) / ext[a]).astype(np.float32)
            q[:, a] = np.clip(np.floor(u * np.float32(1024.0)), 0, 1023)
    want = np.array([interleave(*map(int, qq)) for qq in q], dtype=np.uint32)
    assert np.array_equal(order, np.argsort(want, kind="stable"))
    assert np.array_equal(codes, want[order])
    if n == 1:
        assert nodes.shape[0] == 1 and nodes[0, 6] == order[0] and nodes[0, 7] == 0xFFFFFFFF
        return
    f = nodes.view(np.float32)
    seen = np.zeros(2 * n - 1, dtype=int)

    def walk(i):
        seen[i] += 1
        if nodes[i, 7] == 0xFFFFFFFF:
            k = i - (n - 1)
            assert nodes[i, 6] == order[k]
            return k, k
        l, r = int(nodes[i, 6]), int(nodes[i, 7])
        for c in (l, r):
            assert np.all(f[i, :3] <= f[c, :3]) and np.all(f[i, 3:6] >= f[c, 3:6])
        assert np.array_equal(f[i, :3], np.minimum(f[l, :3], f[r, :3]))
        assert np.array_equal(f[i, 3:6], np.maximum(f[l, 3:6], f[r, 3:6]))
        a0, a1 = walk(l)
        b0, b1 = walk(r)
        assert a1 + 1 == b0  # children cover adjacent ranges of the Morton order
        return a0, b1

    assert walk(0) == (0, n - 1)
    assert np.all(seen == 1)


# ---- bricking ---------------------------------------------------------------------------------
EMPTY = 0xE0000000   # brick map entries at or above this are empty bricks
@pytest.mark.parametrize("dims,fmt", [((32, 32, 32), 0), ((50, 40, 30), 0), ((19, 8, 11), 1)])
def test_brick_roundtrip_and_popcount(dims, fmt):
    idx = scenes.gen_volume(dims, 31, 0.2)
    pal = scenes.gen_palette(31, opaque=(fmt == 0))
    vol = scenes.Volume(idx, pal) if fmt == 0 else scenes.Volume(scenes.expand_palette(idx, pal), None)
    bmap, masks, payload = O.brick(vol)
    dx, dy, dz = dims
    bdx, bdy, bdz = (dx + 7) // 8, (dy + 7) // 8, (dz + 7) // 8
    rm = [np.array([O.lib().orc_nearest_texel(i, d) for i in range(d)]) for d in (dx, dy, dz)]
    texel = vol.data[np.ix_(rm[2], rm[1], rm[0])]
    occ = (texel != 0) if fmt == 0 else ((texel >> 24) != 0)
    bpv = 1 if fmt == 0 else 4
    pay = payload.view(np.uint8 if fmt == 0 else np.uint32)
    recon_occ = np.zeros((bdz * 8, bdy * 8, bdx * 8), dtype=bool)
    recon = np.zeros((bdz * 8, bdy * 8, bdx * 8), dtype=vol.data.dtype)
    for b, e in enumerate(bmap):
        bx, by, bz = b % bdx, (b // bdx) % bdy, b // (bdx * bdy)
        if e >= EMPTY:
            continue
        m = masks[e]
        for z in range(8):
            bits = int(m[z])
            for y in range(8):
                for x in range(8):
                    recon_occ[bz * 8 + z, by * 8 + y, bx * 8 + x] = (bits >> (x + 8 * y)) & 1
        blk = pay[e * 512:(e + 1) * 512].reshape(8, 8, 8)
        recon[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = blk
    assert np.array_equal(recon_occ[:dz, :dy, :dx], occ)
    assert np.array_equal(recon[:dz, :dy, :dx][occ], texel[occ])
    assert int(sum(bin(int(w)).count("1") for w in masks.reshape(-1))) == int(occ.sum())
    # a brick is stored iff it holds an occupied voxel
    full = np.zeros((bdz * 8, bdy * 8, bdx * 8), dtype=bool)
    full[:dz, :dy, :dx] = occ
    blocks = full.reshape(bdz, 8, bdy, 8, bdx, 8).any(axis=(1, 3, 5)).reshape(-1)
    assert np.array_equal(bmap < EMPTY, blocks)
    _ = bpv


def _radius_brute(nonempty):
    """Empty-space radius by brute force: L-inf brick distance to the nearest non-empty brick
    minus 1, capped at 31 (32 - 1); 0 for non-empty bricks."""
    nz = np.argwhere(nonempty)
    out = np.zeros(nonempty.shape, dtype=np.int64)
    for b in np.argwhere(~nonempty):
        d = 32 if len(nz) == 0 else min(32, int(np.abs(nz - b).max(axis=1).min()))
        out[tuple(b)] = d - 1
    return out


@pytest.mark.parametrize("dims,fill,boxes", [((64, 64, 64), 0.03, None), ((96, 40, 72), 0.10, None),
                                             ((48, 48, 48), 0.0, [[20, 21, 22, 21, 22, 23]]),
                                             ((40, 24, 16), 0.0, None)])
def test_brick_radius_matches_brute_force(dims, fill, boxes):
    idx = scenes.gen_volume(dims, 77, fill, boxes=boxes) if fill > 0 or boxes else np.zeros(dims[::-1], np.uint8)
    vol = scenes.Volume(idx, scenes.gen_palette(77))
    bmap, _, _ = O.brick(vol)
    bd = [(d + 7) // 8 for d in dims]
    m = bmap.reshape(bd[2], bd[1], bd[0])
    nonempty = m < EMPTY
    rad = np.where(nonempty, 0, ((m >> 24) & 31).astype(np.int64))
    want = _radius_brute(nonempty)
    assert np.array_equal(rad, want)
    if fill == 0.03:
        assert rad.max() >= 2  # sparse volumes exercise cubes of more than one brick


def _cube_brute(nonempty, octant, cap=8):
    """Side of the largest cube of empty bricks with brick b as its corner, extending towards the
    octant (bit a set: towards -axis a), capped; bricks outside the grid count as empty."""
    bz, by, bx = nonempty.shape
    s = [-1 if (octant >> a) & 1 else 1 for a in range(3)]   # x, y, z signs
    out = np.zeros(nonempty.shape, dtype=np.int64)
    for z, y, x in np.argwhere(~nonempty):
        k = 0
        while k < cap:
            # the cube of side k + 1 is empty iff its bricks inside the grid are all empty
            xs = sorted((x, x + s[0] * k)); ys = sorted((y, y + s[1] * k)); zs = sorted((z, z + s[2] * k))
            sub = nonempty[max(zs[0], 0):min(zs[1], bz - 1) + 1, max(ys[0], 0):min(ys[1], by - 1) + 1,
                           max(xs[0], 0):min(xs[1], bx - 1) + 1]
            if sub.any():
                break
            k += 1
        out[z, y, x] = k
    return out


@pytest.mark.parametrize("dims,fill,boxes", [((64, 64, 64), 0.03, None), ((96, 40, 72), 0.10, None),
                                             ((48, 48, 48), 0.0, [[20, 21, 22, 21, 22, 23]]),
                                             ((40, 24, 16), 0.0, None)])
def test_brick_octant_cubes_match_brute_force(dims, fill, boxes):
    idx = scenes.gen_volume(dims, 78, fill, boxes=boxes) if fill > 0 or boxes else np.zeros(dims[::-1], np.uint8)
    vol = scenes.Volume(idx, scenes.gen_palette(78))
    bmap, _, _ = O.brick(vol)
    bd = [(d + 7) // 8 for d in dims]
    m = bmap.reshape(bd[2], bd[1], bd[0]).astype(np.int64)
    nonempty = m < EMPTY
    seen = set()
    for o in range(8):
        got = np.where(nonempty, 0, ((m >> (3 * o)) & 7) + 1)
        want = _cube_brute(nonempty, o)
        assert np.array_equal(got, want), o
        seen.update(np.unique(got).tolist())
    if fill == 0.03:
        assert max(seen) == 8 and len(seen) >= 4   # cubes of several sizes, up to the cap


@pytest.mark.parametrize("dims,fill", [((32, 32, 32), 0.2), ((50, 40, 30), 0.35), ((19, 8, 11), 0.05)])
def test_block_masks_match_brute_force(dims, fill):
    idx = scenes.gen_volume(dims, 33, fill)
    vol = scenes.Volume(idx, scenes.gen_palette(33))
    blocks = O.brick_blocks(vol)
    dx, dy, dz = dims
    bd = [(d + 7) // 8 for d in dims]
    occ = np.zeros((bd[2] * 8, bd[1] * 8, bd[0] * 8), dtype=bool)
    rm = [np.array([O.lib().orc_nearest_texel(i, d) for i in range(d)]) for d in (dx, dy, dz)]
    occ[:dz, :dy, :dx] = vol.data[np.ix_(rm[2], rm[1], rm[0])] != 0
    blk = occ.reshape(bd[2], 4, 2, bd[1], 4, 2, bd[0], 4, 2).any(axis=(2, 5, 8))   # (bz, z4, by, y4, bx, x4)
    for b, word in enumerate(blocks):
        bx, by, bz = b % bd[0], (b // bd[0]) % bd[1], b // (bd[0] * bd[1])
        want = 0
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    if blk[bz, z, by, y, bx, x]:
                        want |= 1 << (x + 4 * y + 16 * z)
        assert int(word) == want, b


# ---- DDA against brute force -----------------------------------------------------------------------
def _first_voxel_brute(occ, p, d):
    """First occupied voxel entered by p + t d (t >= 0), float64; also the gap to the runner-up."""
    dz, dy, dx = occ.shape
    best = []
    for z, y, x in zip(*np.nonzero(occ)):
        lo = np.array([x, y, z], dtype=np.float64)
        hi = lo + 1.0
        tn, tf = -np.inf, np.inf
        ok = True
        for a in range(3):
            if d[a] == 0:
                if p[a] < lo[a] or p[a] > hi[a]:
                    ok = False
                continue
            t1, t2 = (lo[a] - p[a]) / d[a], (hi[a] - p[a]) / d[a]
            tn, tf = max(tn, min(t1, t2)), min(tf, max(t1, t2))
        if ok and tn <= tf and tf >= 0:
            best.append((max(tn, 0.0), (x, y, z)))
    best.sort()
    if not best:
        return None, np.inf
    gap = best[1][0] - best[0][0] if len(best) > 1 else np.inf
    return best[0][1], gap


@pytest.mark.parametrize("level", [0, 1])
def test_dda_first_hit_matches_brute_force(level):
    rng = np.random.default_rng(40 + level)
    checked = 0
    for trial in range(60):
        n = int(rng.choice([2, 3, 5, 8, 13, 16]))
        occ = rng.random((n, n, n)) < rng.uniform(0.02, 0.3)
        idx = (occ * rng.integers(1, 256, size=occ.shape)).astype(np.uint8)
        vol = scenes.Volume(idx, scenes.gen_palette(1))
        occ = idx != 0
        for _ in range(20):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            if level == 0:  # enter through a face, like a primary ray (mpos on the face)
                a = int(rng.integers(0, 3))
                p = rng.uniform(0.05, n - 0.05, size=3)
                p[a] = 0.0 if d[a] > 0 else float(n)
                v0 = np.floor(np.minimum(p, n - 1)).astype(int)
            else:  # start inside, in a given voxel
                p = rng.uniform(0.0, n, size=3)
                v0 = np.floor(p).astype(int)
            p32 = p.astype(np.float32)
            hit, v, steps, mask, _, _ = O.dda(vol, level, p32, v0, d.astype(np.float32))
            want, gap = _first_voxel_brute(occ, p32.astype(np.float64), d.astype(np.float32).astype(np.float64))
            if level == 1 and occ[tuple(v0[::-1])]:
                assert hit and v == tuple(v0)
                continue
            if gap < 1e-4:
                continue  # the ray passes within rounding distance of a voxel edge: either answer is right
            if want is None:
                assert not hit
            else:
                assert hit and v == tuple(int(c) for c in want), (n, p, d, v, want)
            checked += 1
    assert checked > 500


def test_secondary_dda_crosses_empty_cubes_like_brute_force():
    """Sparse multi-brick volumes (a few small boxes): the secondary walk crosses cubes of
    empty bricks (radius >= 1) in one step and must still find the first occupied voxel."""
    rng = np.random.default_rng(91)
    checked = 0
    big = 0
    for trial in range(12):
        n = int(rng.choice([40, 64, 96]))
        occ = np.zeros((n, n, n), dtype=bool)
        for _ in range(int(rng.integers(2, 7))):
            o = rng.integers(0, n - 4, size=3)
            sz = rng.integers(1, 5, size=3)
            occ[o[2]:o[2] + sz[2], o[1]:o[1] + sz[1], o[0]:o[0] + sz[0]] = True
        idx = (occ * rng.integers(1, 256, size=occ.shape)).astype(np.uint8)
        vol = scenes.Volume(idx, scenes.gen_palette(2))
        bmap, _, _ = O.brick(vol)
        big += int(((bmap >= EMPTY) & (((bmap >> 24) & 31) >= 2)).any())
        targets = np.argwhere(occ)
        for _ in range(40):
            p = rng.uniform(0.0, n, size=3)
            if rng.random() < 0.7:  # aim near an occupied voxel so that most rays hit
                t = targets[rng.integers(len(targets))][::-1] + 0.5
                d = t - p + rng.normal(scale=1.5, size=3)
            else:
                d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            v0 = np.floor(p).astype(int)
            if occ[tuple(v0[::-1])]:
                continue
            p32 = p.astype(np.float32)
            hit, v, _, _, _, _ = O.dda(vol, 1, p32, v0, d.astype(np.float32))
            want, gap = _first_voxel_brute(occ, p32.astype(np.float64), d.astype(np.float32).astype(np.float64))
            if gap < 1e-4:
                continue
            if want is None:
                assert not hit
            else:
                assert hit and v == tuple(int(c) for c in want), (n, p, d, v, want)
            checked += 1
    assert checked > 300 and big >= 6


def test_dda_axis_aligned_reference_quirk():
    """mdir with zero components: 0 * inf = NaN side distances stop the walk after one step
    (trace.frag:84), so only the entry voxel and its +x neighbour are tested."""
    n = 8
    idx = np.zeros((n, n, n), dtype=np.uint8)
    idx[3, 4, 5] = 7  # z=3, y=4, x=5
    vol = scenes.Volume(idx, scenes.gen_palette(1))
    p = np.array([0.0, 4.5, 3.5], dtype=np.float32)
    hit, v, steps, mask, _, _ = O.dda(vol, 0, p, [0, 4, 3], [1.0, 0.0, 0.0])
    assert not hit  # the voxel at x=5 is never reached
    assert steps == 3 * n  # the loop burns max_steps
    p = np.array([4.0, 4.5, 3.5], dtype=np.float32)
    hit, v, steps, mask, _, _ = O.dda(vol, 0, p, [4, 4, 3], [1.0, 0.0, 0.0])
    assert hit and v == (5, 4, 3) and steps == 1
    # the secondary DDA has no such quirk: a long axis-aligned walk reaches the voxel
    hit, v, steps, mask, _, _ = O.dda(vol, 1, np.array([0.0, 4.5, 3.5], np.float32), [0, 4, 3], [1.0, 0.0, 0.0])
    assert hit and v == (5, 4, 3)


def test_dda_mirror_symmetry():
    rng = np.random.default_rng(7)
    n = 12
    occ = rng.random((n, n, n)) < 0.08
    idx = (occ * 9).astype(np.uint8)
    vol = scenes.Volume(idx, scenes.gen_palette(1))
    vol_m = scenes.Volume(np.ascontiguousarray(idx[:, :, ::-1]), scenes.gen_palette(1))
    agree = 0
    for _ in range(300):
        d = rng.normal(size=3).astype(np.float32)
        p = rng.uniform(0.1, n - 0.1, size=3).astype(np.float32)
        p[0] = 0.0 if d[0] > 0 else float(n)
        v0 = np.floor(np.minimum(p, n - 1)).astype(int)
        pm = p.copy()
        pm[0] = np.float32(n) - p[0]
        dm = d.copy()
        dm[0] = -d[0]
        v0m = np.floor(np.minimum(pm, n - 1)).astype(int)
        h1, v1, *_ = O.dda(vol, 0, p, v0, d)
        h2, v2, *_ = O.dda(vol_m, 0, pm, v0m, dm)
        if h1 and h2 and v2 == (n - 1 - v1[0], v1[1], v1[2]):
            agree += 1
        elif h1 == h2 == False:  # noqa: E712
            agree += 1
    assert agree >= 290  # mirror images agree except rounding-level ties


# ---- whole-frame closed forms ---------------------------------------------------------------------
def _plane_scene(mode, spp, max_bounces, sky, sun_rad, albedo_code=255):
    n = 16
    idx = np.zeros((n, n, n), dtype=np.uint8)
    idx[:, 0, :] = 1  # one-voxel-thick floor at y = 0 (model space), a convex slab
    pal = np.zeros(256, dtype=np.uint32)
    pal[1] = albedo_code | (albedo_code << 8) | (albedo_code << 16) | (255 << 24)
    vol = scenes.Volume(idx, pal)
    rec = scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), float(n)), 0)[None, :]
    P = scenes.perspective(scenes.FOV, 1.0)
    C = scenes.look_at((2.0, 10.0, 11.0), (0.0, -7.0, 0.0))
    prm = scenes.RenderParams(48, 48, viewport_square=0, mode=mode, spp=spp, max_bounces=max_bounces, jitter=1,
                              seed=3, sun_dir=tuple(np.array([0.3, -0.9, 0.2]) / np.linalg.norm([0.3, -0.9, 0.2])),
                              sun_radiance=(sun_rad,) * 3, sky=(sky,) * 3)
    return scenes.Scene([vol], rec, P, C, prm)


def test_white_furnace_slab():
    """Albedo 1, sky 1, no sun: every path escapes a convex slab after one bounce, so radiance is 1."""
    sc = _plane_scene(mode=2, spp=4, max_bounces=3, sky=1.0, sun_rad=0.0)
    hits, rad, _ = O.render(sc, use_bvh=True)
    assert (hits["inst"] != 0xFFFFFFFF).mean() > 0.2
    assert np.allclose(rad[..., :3], 1.0, rtol=0, atol=1e-12)


def test_lambertian_slab_under_sun():
    """Mode 1 (no bounces, no sky): unshadowed slab pixels are rho * E * cos / pi exactly."""
    sc = _plane_scene(mode=1, spp=1, max_bounces=0, sky=0.0, sun_rad=2.0, albedo_code=200)
    sc.params.jitter = 0
    hits, rad, _ = O.render(sc, use_bvh=True)
    c = 200 / 255.0
    rho = ((c + 0.055) / 1.055) ** 2.4
    sun = np.array(sc.params.sun_dir, dtype=np.float32).astype(np.float64)
    normals = {0: [-1, 0, 0], 1: [1, 0, 0], 2: [0, -1, 0], 3: [0, 1, 0], 4: [0, 0, -1], 5: [0, 0, 1]}
    checked = 0
    for y in range(48):
        for x in range(48):
            h = hits[y, x]
            if h["inst"] == 0xFFFFFFFF:
                assert np.all(rad[y, x, :3] == 0.0)
                continue
            cos = float(np.dot(normals[int(h["face"])], sun))
            want = rho * 2.0 * cos / np.pi if (cos > 0 and not (h["flags"] & 2)) else 0.0
            assert abs(rad[y, x, 0] - want) <= 1e-6 * max(want, 1e-6)
            checked += 1
    assert checked > 100


def test_camera_inside_box_sees_nothing():
    sc = scenes.single_volume_scene(16, 0.3, 5, 32, 24, mode=0)
    sc.C = scenes.look_at((0.0, 0.0, 0.0), (1.0, 0.2, 0.3))
    hits, _, _ = O.render(sc)
    assert (hits["inst"] == 0xFFFFFFFF).all()


def test_lbvh_and_brute_force_frames_agree():
    sc = scenes.multi_volume_scene(count=5, n=32, world=256, seed=9, width=64, height=40, spp=2, max_bounces=2)
    h0, r0, s0 = O.render(sc, use_bvh=False)
    h1, r1, s1 = O.render(sc, use_bvh=True)
    for f in ("inst", "voxel", "steps", "face", "pal", "mask", "flags"):
        assert np.array_equal(h0[f], h1[f])
    assert np.array_equal(r0, r1)
    assert s1["nodes"] > 0 and s0["nodes"] == 0
    assert s0["rays"] == s1["rays"] and s0["hits"] == s1["hits"]


# ---- bounce sampling (cosine-weighted hemisphere) --------------------------------------------
# The path tracer's bounce rule has no counterpart in the reference (trace.frag:79 has no
# shading), so it is pinned by the law it must follow: for a cosine-weighted direction w about
# n, c = n.w has density 2c on [0, 1], i.e. P(c <= x) = x^2 and E[c] = 2/3, and the azimuth
# about n is uniform.  The same checks are run on two plausible wrong samplers to show they
# reject them.
_NORMALS = [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (-1.0, 0.0, 0.0), (1.0, 2.0, -3.0)]


def _frame(n):
    n = np.asarray(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t = np.cross(n, a)
    t /= np.linalg.norm(t)
    return n, t, np.cross(n, t)


def _cosine_law_stats(w, n):
    from scipy import stats
    n, t, b = _frame(n)
    w = w.astype(np.float64)
    c = w @ n
    phi = np.arctan2(w @ b, w @ t)
    ks_cos = stats.kstest(np.clip(c, 0.0, 1.0), lambda x: np.clip(x, 0.0, 1.0) ** 2)
    ks_phi = stats.kstest(phi, stats.uniform(loc=-np.pi, scale=2 * np.pi).cdf)
    z_mean = (c.mean() - 2.0 / 3.0) / np.sqrt((0.5 - 4.0 / 9.0) / len(c))   # Var[c] = 1/2 - 4/9
    return c, ks_cos.pvalue, ks_phi.pvalue, z_mean


@pytest.mark.parametrize("normal", _NORMALS)
def test_cosine_dir_follows_the_cosine_law(normal):
    n = np.asarray(normal, dtype=np.float32)
    n = (n / np.float32(np.linalg.norm(n))).astype(np.float32)
    w = O.cosine_dir_samples(11, n, 200_000)
    assert np.all(np.isfinite(w))
    assert np.allclose(np.linalg.norm(w.astype(np.float64), axis=1), 1.0, atol=1e-6)
    c, p_cos, p_phi, z = _cosine_law_stats(w, n)
    assert np.all(c > 0.0)                   # upper hemisphere only
    assert p_cos > 1e-3, f"n.w is not distributed as P(c <= x) = x^2 (KS p = {p_cos:.2e})"
    assert p_phi > 1e-3, f"azimuth is not uniform (KS p = {p_phi:.2e})"
    assert abs(z) < 5.0, f"mean cos {c.mean():.5f} is {z:.1f} sigma from 2/3"


def test_cosine_law_checks_reject_wrong_samplers():
    """The checks above have the power to catch a wrong sampler: uniform-hemisphere directions
    (mean cos 1/2) and an un-normalised n + (point in the unit ball) both fail them."""
    rng = np.random.default_rng(5)
    n = np.array([0.0, 0.0, 1.0])
    v = rng.normal(size=(200_000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    uni = np.where((v @ n)[:, None] < 0, -v, v)              # uniform over the hemisphere
    _, p_cos, _, z = _cosine_law_stats(uni, n)
    assert p_cos < 1e-10 and abs(z) > 50
    ball = v * rng.random((len(v), 1)) ** (1.0 / 3.0)        # uniform in the unit ball
    a = n + ball
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    _, p_cos, _, z = _cosine_law_stats(a, n)
    assert p_cos < 1e-10 and abs(z) > 10


def test_cosine_dir_is_deterministic_per_stream():
    n = np.array([0.0, 1.0, 0.0], dtype=np.float32)
    a = O.cosine_dir_samples(3, n, 1000)
    b = O.cosine_dir_samples(3, n, 1000)
    c = O.cosine_dir_samples(4, n, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


# ---- sRGB encode (B8G8R8A8_SRGB store, swapchain.c:88) ----------------------------------------
def _srgb_oetf(x):
    """The published sRGB encoding (IEC 61966-2-1): linear -> non-linear in [0, 1]."""
    x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 1.0)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)


def test_srgb_encode_round_trips_every_code():
    L = O.lib()
    dec = np.array([L.orc_srgb_decode(k) for k in range(256)], dtype=np.float32)
    assert dec[0] == 0.0 and dec[255] == 1.0 and np.all(np.diff(dec) > 0)
    for k in range(256):
        assert L.orc_srgb_encode(float(dec[k])) == k


def test_srgb_encode_matches_the_published_oetf():
    L = O.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.random(60_000), rng.random(20_000) * 0.01, np.linspace(0.0, 1.0, 4001)]).astype(np.float32)
    want_f = 255.0 * _srgb_oetf(xs)
    clear = np.abs(want_f - np.floor(want_f) - 0.5) > 1e-4       # not on a rounding boundary
    want = np.floor(want_f + 0.5).astype(np.int64)
    got = np.array([L.orc_srgb_encode(float(x)) for x in xs], dtype=np.int64)
    assert clear.mean() > 0.99
    assert np.array_equal(got[clear], want[clear])
    assert np.all(np.diff(got[np.argsort(xs, kind="stable")]) >= 0)   # monotone
    assert L.orc_srgb_encode(-0.25) == 0 and L.orc_srgb_encode(3.0) == 255


@pytest.mark.parametrize("n,fill", [(32, 0.1), (40, 0.3), (24, 0.45)])
def test_secondary_walk_on_noise_volumes_matches_brute_force(n, fill):
    """Noise volumes like the workloads' (bricks with empty 4^3 octants and 2^3 blocks next to
    occupied ones): the secondary walk's first hit equals the nearest occupied voxel."""
    rng = np.random.default_rng(int(fill * 100) + n)
    idx = scenes.gen_volume((n, n, n), 5 + n, fill)
    vol = scenes.Volume(idx, scenes.gen_palette(3))
    occ = idx != 0
    free = np.argwhere(~occ)
    checked = hits = 0
    for _ in range(400):
        z, y, x = free[rng.integers(len(free))]
        p = np.array([x, y, z], dtype=np.float64) + rng.uniform(0.0, 1.0, size=3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        v0 = np.floor(p).astype(int)
        p32 = p.astype(np.float32)
        if occ[tuple(np.floor(p32.astype(np.float64)).astype(int)[::-1])] or np.any(np.floor(p32) != v0):
            continue
        hit, v, steps, _, _, q = O.dda(vol, 1, p32, v0, d.astype(np.float32))
        want, gap = _first_voxel_brute(occ, p32.astype(np.float64), d.astype(np.float32).astype(np.float64))
        if gap < 1e-4:
            continue
        if want is None:
            assert not hit
        else:
            assert hit and v == tuple(int(c) for c in want), (p, d, v, want)
            # the walk ends on the hit voxel's boundary
            assert np.all(np.asarray(q) >= np.asarray(v) - 1e-3) and np.all(np.asarray(q) <= np.asarray(v) + 1 + 1e-3)
            hits += 1
        checked += 1
    assert checked > 300 and hits > 100
