"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same seeded inputs.

Bar: hit buffers (instance, voxel, steps, face, palette index, step mask, flags) bit-exact;
mode-0 colour bit-exact (FP32 both sides); modes 1/2 FP32 radiance within 1e-5 relative per
pixel of the FP64 oracle (max and mean printed); traversal counters equal to the oracle's
LBVH replay.
"""
import copy
import sys

import numpy as np
import pytest

import oracle_ffi as O
from parity_util import assert_hits_equal, gpu_render, radiance_error
from vtrace_amd import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_config1_primary_colour_bit_exact():
    sc = scenes.config1()
    p = copy.copy(sc.params)
    p.mode = 0
    _say("case: config1 mode0 640x480")
    g_hits, g_rad, _ = gpu_render(sc, params=p)
    o_hits, o_rad, _ = O.render(sc, params=p)
    assert_hits_equal(g_hits, o_hits, "config1 mode0")
    assert np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32)), "mode-0 colour must be bit-exact"
    assert (g_hits["inst"] != 0xFFFFFFFF).mean() > 0.05


def test_config1_shadow_ray():
    sc = scenes.config1()
    _say("case: config1 mode1 640x480 (primary + shadow)")
    g_hits, g_rad, g_st = gpu_render(sc, stats=True)
    o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
    assert_hits_equal(g_hits, o_hits, "config1 mode1")
    mx, mean = radiance_error(g_rad, o_rad)
    _say(f"config1 radiance rel err max={mx:.3e} mean={mean:.3e}")
    assert mx <= REL_TOL
    assert g_st == o_st, (g_st, o_st)


@pytest.mark.parametrize("square", [0, 1])
def test_small_path_tracing_scene(square):
    sc = scenes.multi_volume_scene(count=6, n=64, world=512, seed=11, width=192, height=108, spp=2, max_bounces=4,
                                   viewport_square=square)
    _say(f"case: path 6x64^3 192x108 2spp 4 bounces square={square}")
    g_hits, g_rad, g_st = gpu_render(sc, stats=True)
    o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
    assert_hits_equal(g_hits, o_hits, "path")
    mx, mean = radiance_error(g_rad, o_rad)
    _say(f"path radiance rel err max={mx:.3e} mean={mean:.3e}")
    assert mx <= REL_TOL
    assert g_st == o_st, (g_st, o_st)


def test_rgba_texture_non_power_of_two_and_alpha():
    """add_texture with RGBA8 texels, 50^3 (the Treasure.vox size), random alpha palette."""
    n = 50
    idx = scenes.gen_volume((n, n, n), 7, 0.25)
    pal = scenes.gen_palette(7, opaque=False)
    rgba = scenes.expand_palette(idx, pal)
    vol = scenes.Volume(rgba, None)
    sc = scenes.single_volume_scene(n, 0.25, 7, 200, 150, mode=0)
    sc.volumes = [vol]
    _say("case: rgba 50^3 200x150 mode0")
    g_hits, g_rad, _ = gpu_render(sc)
    o_hits, o_rad, _ = O.render(sc)
    assert_hits_equal(g_hits, o_hits, "rgba")
    assert np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32))


def test_reduced_config3_sparse():
    sc = scenes.config3(width=320, height=180, n=128)
    _say("case: config3-style 128^3 5% 320x180 1spp 2 bounces")
    g_hits, g_rad, g_st = gpu_render(sc, stats=True)
    o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
    assert_hits_equal(g_hits, o_hits, "c3")
    mx, mean = radiance_error(g_rad, o_rad)
    _say(f"c3 radiance rel err max={mx:.3e} mean={mean:.3e}")
    assert mx <= REL_TOL
    assert g_st == o_st


def test_edge_camera_inside_box_and_axis_aligned():
    """Camera inside the proxy (invisible, trace Q8) and an axis-aligned view (NaN side quirk)."""
    n = 32
    vol = scenes.Volume(scenes.gen_volume((n, n, n), 5, 0.2), scenes.gen_palette(5))
    rec = scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), float(n)), 0)[None, :]
    p = scenes.RenderParams(64, 48, mode=0)
    for eye, target in (((0.0, 0.0, 0.0), (1.0, 0.3, 0.2)), ((0.0, 0.0, 3.0 * n), (0.0, 0.0, 0.0)),
                        ((0.5, 0.5, 3.0 * n), (0.5, 0.5, 0.0))):
        sc = scenes.Scene([vol], rec, scenes.perspective(scenes.FOV, 64 / 48), scenes.look_at(eye, target), p)
        _say(f"case: edge eye={eye}")
        g_hits, g_rad, _ = gpu_render(sc)
        o_hits, o_rad, _ = O.render(sc)
        assert_hits_equal(g_hits, o_hits, f"edge {eye}")
        assert np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32))
    # camera inside: nothing visible
    sc = scenes.Scene([vol], rec, scenes.perspective(scenes.FOV, 64 / 48), scenes.look_at((0, 0, 0), (1, 0.3, 0.2)), p)
    g_hits, _, _ = gpu_render(sc)
    assert (g_hits["inst"] == 0xFFFFFFFF).all()


def test_diagonal_ties_and_partial_bricks():
    """Square frame with an axis-aligned camera over x = y: the diagonal pixels' rays have equal x
    and y direction components and entry coordinates, so their DDA steps x and y together (side
    ties, trace.frag:80-83); 44^3 ends in partial bricks.  Mode 1 adds sun rays along (1, 1, 1)
    (equal components again, the two-level walk's ties).  On the oracle these views end 10-32
    walks with a multi-axis step."""
    n = 44
    vol = scenes.Volume(scenes.gen_volume((n, n, n), 9, 0.15), scenes.gen_palette(9))
    rec = scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), float(n)), 0)[None, :]
    s3 = float(np.float32(1.0) / np.sqrt(np.float32(3.0)))
    for mode in (0, 1):
        p = scenes.RenderParams(256, 256, viewport_square=0, mode=mode, sun_dir=(s3, s3, s3))
        for eye in ((0.37, 0.37, 0.9 * n), (-0.25 * n, -0.25 * n, 0.8 * n), (2.0, 2.0, 0.7 * n)):
            sc = scenes.Scene([vol], rec, scenes.perspective(scenes.FOV, 1.0),
                              scenes.look_at(eye, (eye[0], eye[1], 0.0)), p)
            _say(f"case: diagonal mode={mode} eye={eye}")
            g_hits, g_rad, g_st = gpu_render(sc, stats=True)
            o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
            assert_hits_equal(g_hits, o_hits, f"diagonal {mode} {eye}")
            hit = o_hits["inst"] != 0xFFFFFFFF
            assert (np.isin(o_hits["mask"], [3, 5, 6, 7]) & hit).sum() >= 5
            if mode == 0:
                assert np.array_equal(g_rad.view(np.uint32), o_rad.view(np.uint32))
            else:
                mx, _ = radiance_error(g_rad, o_rad)
                assert mx <= REL_TOL
            assert g_st == o_st, (g_st, o_st)


def test_invalid_texture_id_and_zero_instances():
    n = 16
    vol = scenes.Volume(scenes.gen_volume((n, n, n), 9, 0.5), scenes.gen_palette(9))
    recs = np.stack([scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), float(n)), 0),
                     scenes.instance_record(scenes.model_matrix((40.0, 0.0, 0.0), float(n)), 7)])
    p = scenes.RenderParams(48, 48, mode=0)
    sc = scenes.Scene([vol], recs, scenes.perspective(scenes.FOV, 1.0), scenes.look_at((20, 10, 60), (20, 0, 0)), p)
    _say("case: invalid texture id")
    g_hits, _, _ = gpu_render(sc)
    o_hits, _, _ = O.render(sc)
    assert_hits_equal(g_hits, o_hits, "invalid tex")
    assert not (g_hits["inst"] == 1).any()


def test_gbuffer_consistent_with_hits():
    """G-buffer of sample 0: albedo = sRGB-decoded palette colour of the hit voxel, normal =
    the hit face's outward axis (translate-scale instances), and the point at `distance` along
    the pixel-centre camera ray lies on that face of the hit voxel; misses give the sky."""
    from vtrace_amd import renderer as R
    sc = scenes.multi_volume_scene(count=4, n=64, world=384, seed=13, width=160, height=96, spp=1, max_bounces=2)
    p = copy.copy(sc.params)
    p.jitter = 0
    sc.params = p
    r = R.reset_renderer()
    R.load_scene(r, sc)
    r.enable_gbuffer(True)
    _say("case: G-buffer 4x64^3 160x96")
    r.render(sc.P, sc.C)
    hits, gb = r.hits(), r.gbuffer()
    r.enable_gbuffer(False)
    hit = hits["flags"] & 1 == 1
    assert hit.mean() > 0.2
    assert np.all(gb[~hit][:, 3] == 0.0) and np.all(gb[~hit][:, 4:] == 0.0)
    assert np.allclose(gb[~hit][:, :3], np.array(p.sky, dtype=np.float32))
    dec = np.array([O.lib().orc_srgb_decode(v) for v in range(256)], dtype=np.float32)
    Pm = np.asarray(sc.P, np.float64).reshape(4, 4).T
    Cm = np.asarray(sc.C, np.float64).reshape(4, 4).T
    invC = np.linalg.inv(Cm)
    Rm, O_ = invC[:3, :3], invC[:3, 3]
    ys, xs = np.nonzero(hit)
    for y, x in list(zip(ys, xs))[::7]:
        h = hits[y, x]
        inst = int(h["inst"])
        rec = np.asarray(sc.records[inst], np.float32).copy()
        vol = sc.volumes[int(rec.view(np.uint32)[15])]
        vx, vy, vz = (int(c) for c in R.hit_voxels(h))
        rgba = int(vol.palette[int(h["pal"])])
        assert np.array_equal(gb[y, x, :3], dec[[rgba & 255, (rgba >> 8) & 255, (rgba >> 16) & 255]])
        face = int(h["face"])
        a, neg = face >> 1, (face & 1) == 0
        n = np.zeros(3, np.float32)
        n[a] = -1.0 if neg else 1.0
        assert np.allclose(gb[y, x, 4:7], n, atol=1e-6) and gb[y, x, 7] == 1.0
        nx = 2.0 * (x + 0.5) / p.width - 1.0
        ny = 2.0 * (y + 0.5) / p.height - 1.0
        D = Rm @ np.array([nx / Pm[0, 0], ny / Pm[1, 1], -1.0])
        X = O_ + gb[y, x, 3] * D / np.linalg.norm(D)
        M = rec.astype(np.float64).reshape(4, 4).T
        M[3, 3] = 1.0
        Xm = np.linalg.solve(M, np.append(X, 1.0))[:3]
        v = (Xm + 0.5) * np.array(vol.dims, np.float64)
        vox = np.array([vx, vy, vz], np.float64)
        want = vox[a] + (0.0 if neg else 1.0)
        assert abs(v[a] - want) < 0.05, (y, x, v, vox, face)
        others = [k for k in range(3) if k != a]
        assert np.all(v[others] > vox[others] - 0.05) and np.all(v[others] < vox[others] + 1.05)


def test_vox_model_renders_like_oracle():
    """A model read from .vox bytes (reference layout and palette) through the plain path."""
    import struct
    from vtrace_amd import vox
    rng = np.random.default_rng(8)
    D = 24
    cells = {tuple(int(c) for c in rng.integers(0, D, 3)) for _ in range(2500)}
    voxels = b"".join(struct.pack("<BBBB", x, y, z, int(rng.integers(1, 256))) for (x, y, z) in cells)
    pal = b"".join(struct.pack("<I", int(c) | (255 << 24)) for c in rng.integers(0, 2**24, 256))
    def chunk(cid, content, kids=b""):
        return cid + struct.pack("<II", len(content), len(kids)) + content + kids
    data = b"VOX " + struct.pack("<I", 150) + chunk(b"MAIN", b"", chunk(b"SIZE", struct.pack("<iii", D, D, D)) +
                                                   chunk(b"XYZI", struct.pack("<I", len(cells)) + voxels) +
                                                   chunk(b"RGBA", pal))
    (vol,) = vox.parse(data)
    sc = scenes.single_volume_scene(D, 0.3, 5, 160, 120, mode=1, name="vox24")
    sc.volumes = [vol]
    _say("case: .vox model 24^3 mode1 160x120")
    g_hits, g_rad, _ = gpu_render(sc)
    o_hits, o_rad, _ = O.render(sc)
    assert_hits_equal(g_hits, o_hits, "vox")
    mx, _ = radiance_error(g_rad, o_rad)
    assert mx <= REL_TOL
    assert (g_hits["inst"] != 0xFFFFFFFF).mean() > 0.05


def test_bgra8_encode_matches_oracle():
    """The swapchain format (B8G8R8A8_SRGB, lib/swapchain.c:88): every pixel's bytes from the
    GPU encode of the frame equal the oracle's encode of its own (bit-identical) colour."""
    from vtrace_amd import renderer as R
    sc = scenes.config1(320, 240)
    p = copy.copy(sc.params)
    p.mode = 0
    sc.params = p
    r = R.reset_renderer()
    R.load_scene(r, sc)
    _say("case: bgra8 encode 320x240 mode0")
    r.render(sc.P, sc.C)
    rad, bgra = r.radiance(), r.bgra8()
    enc = np.vectorize(lambda v: O.lib().orc_srgb_encode(float(v)), otypes=[np.uint32])
    b, g, rr = enc(rad[..., 2]), enc(rad[..., 1]), enc(rad[..., 0])
    a = np.clip(np.floor(rad[..., 3].astype(np.float32) * np.float32(255.0) + np.float32(0.5)), 0, 255).astype(np.uint32)
    want = b | (g << 8) | (rr << 16) | (a << 24)
    assert np.array_equal(bgra, want)
    assert len(np.unique(bgra)) > 50


@pytest.mark.parametrize("spp,chunks,gbuffer", [(1, 3, False), (2, 2, False), (3, 3, False), (4, 4, False),
                                                (5, 4, False), (4, 2, False), (4, 4, True), (5, 3, True),
                                                (1, 4, True)])
def test_split_frames_equal_one_stream(spp, chunks, gbuffer):
    """Chunks of tile rows on several streams give the one-stream frame bit for bit (hits,
    radiance, G-buffer, counters): a chunk traces every sample of its pixels, summed in sample
    order; the 100-row frame has 13 rows of tiles (the last one partial), so the chunks are
    uneven."""
    from vtrace_amd import renderer as R
    sc = scenes.multi_volume_scene(count=5, n=64, world=512, seed=21, width=160, height=100, spp=spp, max_bounces=3)
    _say(f"case: split spp={spp} chunks={chunks} gbuffer={gbuffer}")
    r = R.reset_renderer()
    R.load_scene(r, sc, stats=True)
    r.enable_gbuffer(gbuffer)
    r.lib.vt_set_frame_split(1)
    r.render(sc.P, sc.C)
    h1, rad1, st1 = r.hits().copy(), r.radiance().copy(), r.stats()
    gb1 = r.gbuffer().copy() if gbuffer else None
    r.lib.vt_set_frame_split(chunks)
    r.render(sc.P, sc.C)
    assert np.array_equal(r.hits(), h1)
    assert np.array_equal(r.radiance().view(np.uint32), rad1.view(np.uint32))
    assert r.stats() == st1
    if gbuffer:
        assert np.array_equal(r.gbuffer().view(np.uint32), gb1.view(np.uint32))


@pytest.mark.parametrize("blocks", [1, 3])
def test_few_workgroups_overflow_queue_segments(blocks):
    """With 1 or 3 persistent workgroups every queue append starts in segment 0 (or 0-2), so the
    segments overflow into the next ones (the paired bounce-queue and sun-list append's
    fallback) and a secondary wave's pool refills from every part in turn: hits, radiance and
    counters equal the default launch's bit for bit and the oracle's (hits exact, radiance
    within tolerance)."""
    from vtrace_amd import renderer as R
    sc = scenes.multi_volume_scene(count=6, n=64, world=512, seed=11, width=192, height=108, spp=2, max_bounces=4)
    _say(f"case: persistent blocks={blocks}")
    r = R.reset_renderer()
    R.load_scene(r, sc, stats=True)
    r.lib.vt_set_frame_split(1)
    r.render(sc.P, sc.C)
    h1, rad1, st1 = r.hits().copy(), r.radiance().copy(), r.stats()
    assert (h1["inst"] != 0xFFFFFFFF).mean() > 0.2   # more hits than one segment holds
    try:
        assert r.lib.vt_set_persistent_blocks(blocks) == 0
        r.render(sc.P, sc.C)
        h2, rad2, st2 = r.hits().copy(), r.radiance().copy(), r.stats()
    finally:
        r.lib.vt_set_persistent_blocks(0)
    assert np.array_equal(h2, h1)
    assert np.array_equal(rad2.view(np.uint32), rad1.view(np.uint32))
    assert st2 == st1
    o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
    assert_hits_equal(h2, o_hits, f"blocks={blocks}")
    mx, _ = radiance_error(rad2, o_rad)
    assert mx <= REL_TOL
    assert st2 == o_st


def _deep_chain_scene():
    """31 instances of 8^3 volumes (6 world units wide) centred at 8 * 2^k units along each
    axis (k = 0..9), plus one at the far corner: their Morton codes are powers of two, so the
    LBVH is a chain about 30 levels deep; path traced at 160x120, 2 spp, <= 3 bounces."""
    vols = [scenes.Volume(scenes.gen_volume((8, 8, 8), 50 + k, 0.5), scenes.gen_palette(50 + k)) for k in range(4)]
    recs = []
    for a in range(3):
        for k in range(10):
            c = [0.0, 0.0, 0.0]
            c[a] = 8.0 * 2 ** k
            recs.append(scenes.instance_record(scenes.model_matrix(tuple(c), 6.0), len(recs) % 4))
    recs.append(scenes.instance_record(scenes.model_matrix((8184.0, 8184.0, 8184.0), 6.0), 0))
    p = scenes.RenderParams(160, 120, viewport_square=0, mode=2, spp=2, max_bounces=3)
    return scenes.Scene(vols, np.stack(recs).astype(np.float32), scenes.perspective(scenes.FOV, 160 / 120),
                        scenes.look_at((-12.0, 10.0, -12.0), (20.0, 0.0, 20.0)), p, name="deep_chain_31")


def _tree_depth(nodes):
    depth, stack = 0, [(0, 1)]
    while stack:
        i, k = stack.pop()
        depth = max(depth, k)
        if nodes[i, 7] != 0xFFFFFFFF:
            stack += [(int(nodes[i, 6]), k + 1), (int(nodes[i, 7]), k + 1)]
    return depth


def test_deep_tree_traversal():
    """A tree deeper than the traversal stack held in LDS (16 entries): primary queries and the
    secondary search use the per-lane stack in private memory.  Hits bit-exact, radiance within
    tolerance and every counter equal to the oracle's."""
    sc = _deep_chain_scene()
    o_nodes, _, _ = O.lbvh(sc.records, len(sc.volumes))   # the device builds the same tree
    assert _tree_depth(o_nodes) > 17
    _say("case: deep LBVH chain, 31 instances, 160x120 2 spp 3 bounces")
    g_hits, g_rad, g_st = gpu_render(sc, stats=True)
    o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
    assert_hits_equal(g_hits, o_hits, "deep chain")
    assert (g_hits["inst"] != 0xFFFFFFFF).mean() > 0.1
    mx, _ = radiance_error(g_rad, o_rad)
    assert mx <= REL_TOL
    assert g_st == o_st, (g_st, o_st)


def test_full_size_benchmark_frame():
    """The benchmark workload at its full size (16 x 256^3 instances, 1920x1080, 4 spp, <= 4
    bounces) on the benchmark's path: frames of two, three (the default) and four chunks of tile
    rows equal the one-stream frame bit for bit, and bands of rows of it match the oracle (hits
    bit-exact, radiance within 1e-5)."""
    from vtrace_amd import renderer as R
    sc = scenes.config2()
    _say("case: full-size benchmark frame 1920x1080 4 spp")
    r = R.reset_renderer()
    R.load_scene(r, sc)
    r.lib.vt_set_frame_split(1)
    r.render(sc.P, sc.C)
    h1, rad1 = r.hits().copy(), r.radiance().copy()
    for chunks in (2, 3, 4):
        r.lib.vt_set_frame_split(chunks)
        r.render(sc.P, sc.C)
        assert np.array_equal(r.hits(), h1)
        assert np.array_equal(r.radiance().view(np.uint32), rad1.view(np.uint32))
    assert (h1["inst"] != 0xFFFFFFFF).mean() > 0.2
    osc = O.OracleScene(sc, use_bvh=True)
    for y0 in (0, 200, 360, 540, 720, 1064):
        o_hits, o_rad, _ = osc.render(rows=(y0, y0 + 16), threads=16)
        assert_hits_equal(h1[y0:y0 + 16], o_hits[y0:y0 + 16], f"rows {y0}")
        mx, _ = radiance_error(rad1[y0:y0 + 16], o_rad[y0:y0 + 16])
        _say(f"rows {y0}-{y0 + 15}: radiance rel err max={mx:.3e}")
        assert mx <= REL_TOL
    osc.close()


def test_full_size_sparse_volume_frame():
    """The sparse 1024^3 workload at its full size (5 % fill, 3840x2160, 1 spp, 2 bounces; a
    128^3 brick map with empty-cube radii up to 31): bands of rows against the oracle."""
    sc = scenes.config3()
    _say("case: full-size 1024^3 3840x2160 frame")
    g_hits, g_rad, _ = gpu_render(sc)
    assert (g_hits["inst"] != 0xFFFFFFFF).mean() > 0.05
    osc = O.OracleScene(sc, use_bvh=True)
    for y0 in (600, 1080, 1500, 2000):
        o_hits, o_rad, _ = osc.render(rows=(y0, y0 + 8), threads=16)
        assert_hits_equal(g_hits[y0:y0 + 8], o_hits[y0:y0 + 8], f"rows {y0}")
        mx, _ = radiance_error(g_rad[y0:y0 + 8], o_rad[y0:y0 + 8])
        _say(f"rows {y0}-{y0 + 7}: radiance rel err max={mx:.3e}")
        assert mx <= REL_TOL
    osc.close()
