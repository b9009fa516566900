"""GPU: the allocation-free frame entry point (vt_render_into) against render_tick.

vt_render_into renders with the library's static scratch array into caller-owned device
buffers (here torch tensors); render_tick renders the same frame with the library's own
buffers.  Two calls in a row, the second on a smaller frame (the scratch keeps the first
call's contents), each bit-equal to render_tick's hits and radiance; a frame whose state does
not fit the scratch is refused.
"""
import copy
import ctypes

import numpy as np
import pytest

from vtrace_amd import renderer as R
from vtrace_amd import scenes
from vtrace_amd.abi import RenderTickInfo

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _into(lib, sc, torch):
    w, h = sc.params.width, sc.params.height
    hits = torch.zeros(w * h * 16, dtype=torch.uint8, device="cuda")
    rad = torch.full((h, w, 4), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    P = np.ascontiguousarray(sc.P, dtype=np.float32)
    Cm = np.ascontiguousarray(sc.C, dtype=np.float32)
    info = RenderTickInfo(P.ctypes.data, Cm.ctypes.data)
    rc = lib.vt_render_into(ctypes.byref(info), hits.data_ptr(), hits.numel(), rad.data_ptr(), rad.numel() * 4)
    return rc, hits.cpu().numpy().view(R.HIT_DTYPE).reshape(h, w), rad.cpu().numpy()


def _plain(r, sc):
    r.set_params(sc.params)
    r.render(sc.P, sc.C)
    return r.hits(), r.radiance()


def _check(r, sc, torch, label):
    want_hits, want_rad = _plain(r, sc)
    rc, hits, rad = _into(r.lib, sc, torch)
    assert rc == 0, r.error()
    for f in ("inst", "voxel", "steps", "face", "pal", "mask", "flags"):
        assert np.array_equal(hits[f], want_hits[f]), f"{label}: hit field {f}"
    assert np.array_equal(rad, want_rad), f"{label}: radiance"
    assert (hits["inst"] != 0xFFFFFFFF).mean() > 0.05


def test_render_into_equals_render_tick_twice_then_smaller(fresh_renderer):
    torch = pytest.importorskip("torch")
    r = fresh_renderer
    sc = scenes.config1()
    R.load_scene(r, sc)
    _say("case: vt_render_into 640x480 1 spp primary + shadow (128^3)")
    _check(r, sc, torch, "call 1")
    small = copy.copy(sc)
    small.params = copy.copy(sc.params)
    small.params.width, small.params.height = 320, 200
    small.params.mode, small.params.spp, small.params.max_bounces, small.params.jitter = 2, 2, 2, 1
    _say("case: vt_render_into 320x200 2 spp 2 bounces after the larger frame")
    _check(r, small, torch, "call 2 (smaller)")
    prim = copy.copy(sc)
    prim.params = copy.copy(sc.params)
    prim.params.width, prim.params.height, prim.params.mode = 200, 120, 0
    _say("case: vt_render_into 200x120 mode 0")
    _check(r, prim, torch, "call 3 (mode 0)")
    R.reset_renderer()


def test_render_into_refuses_frames_beyond_the_scratch(fresh_renderer):
    torch = pytest.importorskip("torch")
    r = fresh_renderer
    sc = scenes.config1()
    R.load_scene(r, sc)
    big = copy.copy(sc)
    big.params = copy.copy(sc.params)
    big.params.width, big.params.height, big.params.mode, big.params.spp = 1920, 1080, 2, 1
    r.set_params(big.params)
    _say("case: vt_render_into 1920x1080 (beyond the scratch) refused")
    rc, _, _ = _into(r.lib, big, torch)
    assert rc == -1 and "128 MiB" in r.error()
    _say("case: vt_render_into output buffer too small refused")
    r.set_params(sc.params)
    P = np.ascontiguousarray(sc.P, dtype=np.float32)
    info = RenderTickInfo(P.ctypes.data, np.ascontiguousarray(sc.C, dtype=np.float32).ctypes.data)
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    assert r.lib.vt_render_into(ctypes.byref(info), buf.data_ptr(), 64, buf.data_ptr(), 64) == -1
    R.reset_renderer()
