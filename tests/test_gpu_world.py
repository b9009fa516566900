"""GPU: the streamed terrain world — batched chunk upload and rendering of paged scenes.

Bar: every volume of a vt_add_textures batch is bit-identical to the same volume uploaded
alone (brick map, dense read-back) and to the oracle's bricking; frames of the streamed
world match the oracle (hits bit-exact, counters equal, radiance within 1e-5 relative of
the FP64 oracle); an unchanged instance list does not rebuild the LBVH.
"""
import math
import sys

import numpy as np
import pytest

import oracle_ffi as O
from parity_util import assert_hits_equal, radiance_error
from vtrace_amd import renderer as R
from vtrace_amd import scenes
from vtrace_amd import world as W

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_batched_upload_equals_single_uploads():
    r = R.reset_renderer()
    gen = W.TerrainGenerator(0)
    pos = [(x, y, z) for x in range(-3, 3) for y in range(-3, 3) for z in (-1, 0, 1)]
    datas = [v.data for v in gen.gen_chunks(pos) if v is not None][:48]
    datas.append(np.zeros((16, 16, 16), np.uint32))                 # nothing occupied
    datas.append(np.full((16, 16, 16), 0xFF203040, np.uint32))      # everything occupied
    _say(f"case: {len(datas)} chunks of 16^3, one at a time then as one batch")
    single = [r.add_texture_rgba(d) for d in datas]
    batch = r.add_textures_rgba(datas)
    assert batch == list(range(single[-1] + 1, single[-1] + 1 + len(datas)))
    for a, b, d in zip(single, batch, datas):
        ia, ib = r.volume_info(a), r.volume_info(b)
        assert ia["nbricks"] == ib["nbricks"] and ia["dims"] == ib["dims"] == (16, 16, 16)
        bm = r.brick_map(b)
        assert np.array_equal(r.brick_map(a), bm)
        assert np.array_equal(r.volume_dense(b), d)
        omap, _, _ = O.brick(scenes.Volume(d, None))
        assert np.array_equal(bm, omap)


def test_batched_upload_of_larger_volumes():
    """Same-size volumes with several bricks per axis and a non-multiple-of-8 size."""
    r = R.reset_renderer()
    vols = []
    for k in range(5):
        idx = scenes.gen_volume((40, 24, 33), 30 + k, 0.05 + 0.1 * k)
        vols.append(scenes.expand_palette(idx, scenes.gen_palette(30 + k)))
    _say("case: 5 RGBA volumes 40x24x33 batched")
    ids = r.add_textures_rgba(vols)
    for tid, v in zip(ids, vols):
        omap, _, _ = O.brick(scenes.Volume(v, None))
        assert np.array_equal(r.brick_map(tid), omap)
        dense = r.volume_dense(tid)
        # empty bricks keep no payload (their transparent texels read back as 0): compare with
        # the same volume uploaded alone, and the occupied texels with the source
        assert np.array_equal(dense, r.volume_dense(r.add_texture_rgba(v)))
        occ = (v >> 24) != 0
        assert np.array_equal(dense[occ], v[occ])


def _camera(eye, w, h):
    P = scenes.perspective(80.0 / 180.0 * 3.1415926, w / h)
    C = scenes.look_at(eye, (0.0, 0.0, 0.0))
    return P, C


def _oracle_scene(r, pager, scene, P, C, params):
    """The same frame for the oracle: volumes in texture-id order, records as the renderer wrote them."""
    by_id = {}
    for p, hit in pager.chunks.items():
        if hit is not None and hit[1] in r.texture_handle_lookup:
            by_id[r.texture_handle_lookup[hit[1]]] = hit[0]
    vols = [by_id[i] for i in range(len(by_id))]
    recs = [scenes.instance_record(m, r.texture_handle_lookup[h]) for m, h in scene if h in r.texture_handle_lookup]
    return scenes.Scene(vols, np.array(recs, np.float32).reshape(-1, 16), P, C, params, name="terrain")


def test_streamed_terrain_frames_match_oracle():
    r = R.reset_renderer()
    q = R.TextureUploadQueue()
    pager = W.WorldPager(0)
    w, h = 128, 96
    params = scenes.RenderParams(w, h, mode=2, spp=1, max_bounces=2, seed=5, sun_dir=(0.3, -0.8, 0.5))
    r.set_params(params, stats=True)
    for step, eye in enumerate(((0.0, 2.0, -24.0), (1.0, 3.0, -14.0))):
        P, C = _camera(eye, w, h)
        scene = W.terrain_scene(eye, pager, q)
        queued = len(q.queue)
        assert r.render_tick(P, C, q, uploads_per_tick=None)   # drains the queue in one batch
        assert len(q.queue) == 0
        r.update_instances(scene)
        builds = r.lbvh_builds()
        r.update_instances(scene)                              # unchanged: tables and tree kept
        assert r.lbvh_builds() == builds
        r.render(P, C)
        hits, rad, st = r.hits(), r.radiance(), r.stats()
        sc = _oracle_scene(r, pager, scene, P, C, params)
        _say(f"case: terrain step {step} eye={eye} chunks={len(sc.volumes)} queued={queued} "
             f"instances={sc.records.shape[0]}")
        assert sc.records.shape[0] > 10
        o_hits, o_rad, o_st = O.render(sc, use_bvh=True)
        assert_hits_equal(hits, o_hits, f"terrain{step}")
        mx, mean = radiance_error(rad, o_rad)
        _say(f"terrain{step} radiance rel err max={mx:.3e} mean={mean:.3e}")
        assert mx <= REL_TOL
        assert st == o_st, (st, o_st)
        assert (hits["inst"] != 0xFFFFFFFF).mean() > 0.05


def test_reference_pacing_one_upload_per_tick():
    """uploads_per_tick=1 (render.rs:246): instances appear as their textures arrive."""
    r = R.reset_renderer()
    q = R.TextureUploadQueue()
    pager = W.WorldPager(0)
    w, h = 64, 48
    r.set_params(scenes.RenderParams(w, h, mode=0), stats=False)
    eye = (0.0, 1.0, -12.0)
    P, C = _camera(eye, w, h)
    scene = W.terrain_scene(eye, pager, q)
    total = len(q.queue)
    seen = []
    for _ in range(5):
        assert r.render_tick(P, C, q)
        r.update_instances(scene)
        seen.append(r.lib.vt_instance_count())
    assert len(q.queue) == total - 5
    assert seen == [1, 2, 3, 4, 5]
