"""GPU: the continuous-batching serving loop.

Each request's result must be the same alone, in a mix, through a replayed hipGraph and
through the plain single-frame path: bit-exact, because a request's samples depend only on
its own scene, camera, seed and sample indices and are summed in sample order.
"""
import sys

import numpy as np
import pytest

from vtrace_amd import renderer as R
from vtrace_amd import scenes, serving

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(r):
    vols = [scenes.Volume(scenes.gen_volume((n, n, n), 70 + i, f), scenes.gen_palette(70 + i))
            for i, (n, f) in enumerate([(32, 0.3), (64, 0.15), (48, 0.25)])]
    ids = [r.upload(v) for v in vols]
    recs = serving.scene_records(ids, [v.dims[0] for v in vols])
    # a two-instance scene too
    two = np.concatenate([scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), 32.0), ids[0])[None, :],
                          scenes.instance_record(scenes.model_matrix((60.0, 5.0, -10.0), 48.0), ids[2])[None, :]])
    recs.append(two)
    return vols, recs


def _requests():
    rng = scenes.Rng(77)
    out = []
    for i, (sc, w, h, spp) in enumerate([(0, 64, 48, 3), (1, 100, 60, 5), (2, 33, 17, 1), (3, 80, 80, 4),
                                         (1, 17, 40, 2)]):
        n = (32, 64, 48, 60)[sc]
        eye, target = scenes.orbit_camera(rng, (0.0, 0.0, 0.0) if sc < 3 else (30.0, 0.0, 0.0), 1.6 * n, 2.0)
        out.append(serving.Request(sc, w, h, spp, scenes.perspective(scenes.FOV, w / h), scenes.look_at(eye, target),
                                   seed=500 + i, max_bounces=3, sun_dir=scenes.sun_direction(rng)))
    return out


def _plain(r, recs, q):
    r.set_instances(recs[q.scene])
    r.set_params(q.params())
    r.render(q.P, q.C)
    return r.radiance(), r.hits()


@pytest.mark.parametrize("policy,pipeline,poll", [(serving.ROUND_ROBIN, 1, False), (serving.HEAD_RIDERS, 2, False),
                                                  (serving.SHORTEST_FIRST, 2, False), (serving.HEAD_RIDERS, 2, True),
                                                  (serving.ROUND_ROBIN, 1, True)])
def test_mixed_batch_equals_plain_render(policy, pipeline, poll):
    r = R.reset_renderer()
    _, recs = _setup(r)
    reqs = _requests()
    loop = serving.ServingLoop(r, max_slots=8, spp_chunk=2, memory_budget=64 << 20, max_work=1 << 15, tile_w=8,
                               max_bounces=3, use_graph=False, policy=policy, pipeline=pipeline, poll=poll)
    for rc in recs:
        loop.add_scene(rc)
    _say(f"case: serving mixed batch of 5 vs plain render, policy {policy} pipeline {pipeline} poll {poll}")
    got = loop.run_all(reqs)
    loop.close()
    for i, q in enumerate(reqs):
        want, _ = _plain(r, recs, q)
        assert np.array_equal(got[i].view(np.uint32), want.view(np.uint32)), f"request {i}"


def test_alone_vs_mixed_and_graph_replays():
    r = R.reset_renderer()
    _, recs = _setup(r)
    reqs = _requests()
    results = {}
    for tag, graph, subset in (("alone", False, [1]), ("mixed", False, None), ("graph", True, None)):
        loop = serving.ServingLoop(r, max_slots=8, spp_chunk=1, memory_budget=64 << 20, max_work=1 << 20, tile_w=16,
                                   max_bounces=3, use_graph=graph, policy=serving.ROUND_ROBIN, pipeline=2)
        for rc in recs:
            loop.add_scene(rc)
        sel = [reqs[i] for i in subset] if subset else reqs
        _say(f"case: serving {tag}")
        out = loop.run_all(sel)
        st = loop.status()
        loop.close()
        results[tag] = out if subset is None else {1: out[0]}
        if graph:
            assert st["iterations"] >= 5  # spp_chunk 1: the graph was replayed for every iteration
    assert np.array_equal(results["alone"][1], results["mixed"][1])
    for i in range(len(reqs)):
        assert np.array_equal(results["mixed"][i], results["graph"][i]), f"request {i}"


@pytest.mark.parametrize("graph,policy", [(False, serving.ROUND_ROBIN), (True, serving.ROUND_ROBIN),
                                          (True, serving.HEAD_RIDERS)])
def test_sub_batches_on_streams_equal_plain(graph, policy):
    """Iterations large enough to be spread over several streams (requests cut into consecutive
    sample pieces of consecutive sub-batches) give every request the plain render bit for bit."""
    r = R.reset_renderer()
    _, recs = _setup(r)
    rng = scenes.Rng(91)
    reqs = []
    for i, (sc, w, h, spp) in enumerate([(1, 320, 240, 6), (3, 256, 200, 5), (0, 300, 180, 3), (2, 160, 120, 7)]):
        n = (32, 64, 48, 60)[sc]
        eye, target = scenes.orbit_camera(rng, (0.0, 0.0, 0.0) if sc < 3 else (30.0, 0.0, 0.0), 1.6 * n, 2.0)
        reqs.append(serving.Request(sc, w, h, spp, scenes.perspective(scenes.FOV, w / h), scenes.look_at(eye, target),
                                    seed=900 + i, max_bounces=3, sun_dir=scenes.sun_direction(rng)))
    loop = serving.ServingLoop(r, max_slots=8, spp_chunk=4, memory_budget=256 << 20, max_work=1 << 20, tile_w=8,
                               max_bounces=3, use_graph=graph, policy=policy, pipeline=2, streams=4)
    for rc in recs:
        loop.add_scene(rc)
    _say(f"case: serving sub-batches on 4 streams, graph {graph}, policy {policy}")
    got = loop.run_all(reqs)
    st = loop.status()
    loop.close()
    assert st["sub_batches"] > st["iterations"], st
    for i, q in enumerate(reqs):
        want, _ = _plain(r, recs, q)
        assert np.array_equal(got[i].view(np.uint32), want.view(np.uint32)), f"request {i}"


def test_budget_admission_queues_and_completes():
    r = R.reset_renderer()
    _, recs = _setup(r)
    reqs = _requests()
    budget = 80 * 80 * 48 + 4096  # room for the largest request, about one at a time
    loop = serving.ServingLoop(r, max_slots=8, spp_chunk=2, memory_budget=budget, max_work=1 << 20, tile_w=8,
                               max_bounces=3, use_graph=False, pipeline=1)
    for rc in recs:
        loop.add_scene(rc)
    ids = [loop.submit(q) for q in reqs]
    first = loop.step()
    st = loop.status()
    assert st["queued"] >= 3 and st["active"] <= 2
    assert st["memory_in_use"] <= budget
    done = {}
    # a finished request holds its buffers until it is read: read every completion
    for _ in range(200):
        for rid in first + loop.step():
            q = reqs[ids.index(rid)]
            done[rid] = loop.read(rid, q.width, q.height)
        first = []
        if len(done) == len(reqs):
            break
    loop.close()
    assert len(done) == len(reqs)
    want, _ = _plain(r, recs, reqs[0])
    assert np.array_equal(done[ids[0]].view(np.uint32), want.view(np.uint32))


def test_poisson_stream_latency_and_autotuner(tmp_path):
    r = R.reset_renderer()
    vols, recs = _setup(r)
    dims = [32, 64, 48, 60]
    reqs = serving.request_stream(12, len(recs), dims, rate_hz=200.0, seed=9, max_bounces=3,
                                  resolutions=((64, 64), (96, 54)), spp_range=(1, 4))
    tune = serving.autotune_tile(r, reqs[:4], recs, cache_path=str(tmp_path / "tune.json"), key="test", iters=1,
                                 max_slots=8, spp_chunk=2, memory_budget=64 << 20, max_work=1 << 20, max_bounces=3)
    assert tune["tile_w"] in serving.TILE_CANDIDATES
    loop = serving.ServingLoop(r, max_slots=8, spp_chunk=2, memory_budget=64 << 20, max_work=1 << 20,
                               tile_w=tune["tile_w"], max_bounces=3, use_graph=True)
    for rc in recs:
        loop.add_scene(rc)
    res = loop.run_stream(reqs)
    loop.close()
    p = serving.percentiles(res["latency_s"])
    _say("stream latency", p)
    assert np.all(res["latency_s"] > 0) and p["p99_ms"] >= p["p50_ms"]
