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
