"""GPU: the serving loop and the denoiser at their own workloads.

* Serving: requests of the serving workload's distribution (resolution from {256x256, 640x480,
  1280x720, 1920x1080}, spp 1..16, up to 4 bounces, over the 32 resident noise volumes of
  64^3..512^3) run through the loop as the benchmark runs it (hipGraph replay, four sub-batch
  streams, head + riders, pipeline 2, polling steps).  Every result equals the plain
  render_tick render bit for bit, and bands of rows of a 1920x1080 request match the oracle
  (hits bit-exact, radiance within 1e-5 relative).
* Denoiser: vt_dn_denoise_frame on a 1920x1080 frame of the path-tracing benchmark scene
  (padded to 1088 rows inside the library) against the reference network, which rounds to
  BF16 at the same points, on tiles around blocks of the frame (corner, bottom edge with the
  padded rows, interior) with a halo wider than the network's receptive field.
"""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as O
from parity_util import assert_hits_equal, radiance_error
from vtrace_amd import denoise as DN
from vtrace_amd import renderer as R
from vtrace_amd import scenes, serving

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5
# The U-Net's dependence radius measured on the CPU is at most 106 pixels (5 levels of two
# 3x3 convolutions plus pooling alignment); tests/test_denoise_ref.py pins that a 144-pixel
# halo reproduces the whole-frame forward exactly.
HALO = 144


def _say(*a):
    print(*a, flush=True)


def test_serving_stream_at_its_workload():
    r = R.reset_renderer()
    vols = serving.resident_volumes()
    dims = [v.dims[0] for v in vols]
    assert 512 in dims and 64 in dims
    ids = [r.upload(v) for v in vols]
    recs = serving.scene_records(ids, dims)
    reqs = serving.request_stream(24, len(recs), dims, rate_hz=1.0, seed=23, max_bounces=4)
    assert any((q.width, q.height) == (1920, 1080) for q in reqs)
    assert max(q.spp for q in reqs) >= 12
    # the 1920x1080 request checked against the oracle: prefer one over a 512^3 volume
    big = [i for i, q in enumerate(reqs) if (q.width, q.height) == (1920, 1080)]
    pick = next((i for i in big if dims[reqs[i].scene] == 512), big[0])
    loop = serving.ServingLoop(r, max_slots=32, spp_chunk=4, memory_budget=2 << 30, max_work=16 << 20, tile_w=4,
                               max_bounces=4, use_graph=True, policy=serving.HEAD_RIDERS, pipeline=2, streams=4,
                               poll=True)
    for rc in recs:
        loop.add_scene(rc)
    _say(f"case: serving 24 requests of the serving workload (largest {max(q.width * q.height for q in reqs)} px, "
         f"spp up to {max(q.spp for q in reqs)})")
    ids_ = {loop.submit(q): i for i, q in enumerate(reqs)}
    out, hits = {}, {}
    while len(out) < len(reqs):
        for rid in loop.step():
            i = ids_[rid]
            q = reqs[i]
            if i == pick:
                out[i], hits[i] = loop.read(rid, q.width, q.height, hits=True)
            else:
                out[i] = loop.read(rid, q.width, q.height)
    st = loop.status()
    loop.close()
    assert st["sub_batches"] >= st["iterations"]
    for i, q in enumerate(reqs):
        r.set_instances(recs[q.scene])
        r.set_params(q.params())
        r.render(q.P, q.C)
        want = r.radiance()
        assert np.array_equal(out[i].view(np.uint32), want.view(np.uint32)), f"request {i} ({q.width}x{q.height})"
        if i == pick:
            assert np.array_equal(hits[i], r.hits())
    q = reqs[pick]
    sc = scenes.Scene([vols[q.scene]], scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0),
                                                                                   float(dims[q.scene])), 0)[None, :],
                      q.P, q.C, q.params(), name="serving_pick")
    _say(f"case: served 1920x1080 request {pick} ({dims[q.scene]}^3, {q.spp} spp) vs the oracle")
    osc = O.OracleScene(sc, use_bvh=True)
    g_hits, g_rad = hits[pick], out[pick]
    assert (g_hits["inst"] != 0xFFFFFFFF).mean() > 0.05
    for y0 in (100, 400, 540, 800, 1064):
        o_hits, o_rad, _ = osc.render(rows=(y0, y0 + 8), threads=16)
        assert_hits_equal(g_hits[y0:y0 + 8], o_hits[y0:y0 + 8], f"rows {y0}")
        mx, mean = radiance_error(g_rad[y0:y0 + 8], o_rad[y0:y0 + 8])
        _say(f"rows {y0}-{y0 + 7}: radiance rel err max={mx:.3e} mean={mean:.3e}")
        assert mx <= REL_TOL
    osc.close()


def _tile(c0, size, total):
    lo = max(0, (c0 - HALO) // 16 * 16)
    hi = min(total, -(-(c0 + size + HALO) // 16) * 16)
    return lo, hi


def test_denoiser_on_a_1080p_frame():
    r = R.reset_renderer()
    sc = scenes.config2(spp=1)
    R.load_scene(r, sc)
    d = DN.Denoiser(r, seed=5)
    rng = np.random.default_rng(8)
    p = d.init_params(8)
    for cin, cout, wo, bo in d.layers:
        p[bo:bo + cout] = rng.normal(0.0, 0.05, cout)
    d.set_params(p)
    r.enable_gbuffer(True)
    r.render(sc.P, sc.C)
    _say("case: vt_dn_denoise_frame 1920x1080 (padded to 1088) vs the reference network on tiles")
    out, ms = d.denoise_frame()
    rad, gb = r.radiance(), r.gbuffer()
    r.enable_gbuffer(False)
    d.close()
    H, W = rad.shape[:2]
    assert (H, W) == (1080, 1920) and np.all(out[..., 3] == 1.0) and np.all(np.isfinite(out))
    f = np.pad(U.features(rad, gb, 1), ((0, 1088 - H), (0, 0), (0, 0)), mode="edge")
    net = U.UNetRef(d.layers, DN.ENC, DN.DEC)
    p64 = p.astype(np.float64)
    for cy, cx in ((0, 0), (1048, 944), (520, 960)):
        y0, y1 = _tile(cy, 32, 1088)
        x0, x1 = _tile(cx, 32, W)
        y, c = net.forward(p64, f[None, y0:y1, x0:x1], bf16=True)
        want = U.radiance(y, c["x0"])[0, cy - y0:cy - y0 + 32, cx - x0:cx - x0 + 32]
        got = out[cy:cy + 32, cx:cx + 32, :3]
        e = float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-30))
        _say(f"block ({cy}, {cx}) in tile [{y0}:{y1}, {x0}:{x1}]: rel err {e:.2e} ({ms:.3f} ms per frame)")
        assert e < 1e-2
