"""GPU: the headless host loop through librender_hip — cursor/last_mouse semantics of
render_tick (lib/entry.c:157-163), the one-frame camera latency of render.rs:240-320, and a short
main.rs loop over entities and streamed terrain."""
import sys

import numpy as np
import pytest

from vtrace_amd import app, scenes
from vtrace_amd import renderer as R

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _entities():
    a = np.zeros((8, 8, 8), np.uint32)
    a[1:7, 1:7, 1:7] = 0xFF30C050
    b = np.zeros((12, 12, 12), np.uint32)
    b[:, :5, :] = 0xFFA04020
    return scenes.Volume(a, None), scenes.Volume(b, None)


def test_cursor_becomes_mouse_and_last_mouse():
    r = R.reset_renderer()
    r.set_params(scenes.RenderParams(32, 24, mode=0))
    ptr = r.lib.get_input_data_pointer()
    P, Cm = scenes.perspective(1.2, 32 / 24), scenes.look_at((0, 0, -5), (0, 0, 0))
    _say("case: cursor sequence over three render_ticks")
    r.lib.vt_set_cursor(100.0, 50.0)
    assert r.render_tick(P, Cm)
    u = ptr.contents
    assert (u.mouse_x, u.mouse_y, u.last_mouse_x, u.last_mouse_y) == (100.0, 50.0, 100.0, 50.0)  # first frame
    r.lib.vt_set_cursor(110.0, 40.0)
    assert r.render_tick(P, Cm)
    assert (u.mouse_x, u.mouse_y, u.last_mouse_x, u.last_mouse_y) == (110.0, 40.0, 100.0, 50.0)
    assert r.render_tick(P, Cm)
    assert (u.mouse_x, u.last_mouse_x) == (110.0, 110.0)
    ptr.contents.keys[0] = 1                     # keys are written by the host directly
    assert app.Input.from_ptr(ptr).key_w == 1


def test_one_frame_camera_latency_on_gpu():
    r = R.reset_renderer()
    q = R.TextureUploadQueue()
    w = app.WorldState(q, _entities())
    w.camera_position = np.array([0.5, -3.0, -12.0], np.float32)
    scene = w.update(0.0, app.Input(), q)
    W_, H_ = 96, 64
    r.set_params(scenes.RenderParams(W_, H_, mode=0))
    t = app.CameraTicker(r)
    assert t.tick(w.camera_position, w.get_camera_direction(), q, uploads_per_tick=None)[0]
    r.update_instances(scene)
    a = np.array([1.0, -2.0, -14.0], np.float32)
    da = np.array([0.0, 0.1, 1.0], np.float32)
    b = np.array([-3.0, 4.0, -9.0], np.float32)
    _say("case: ticks at camera A then B; the second frame is rendered from A")
    assert t.tick(a, da, q)[0]
    assert t.tick(b, da, q)[0]
    hits_tick = r.hits().copy()
    rad_tick = r.radiance().copy()
    r.render(app.CameraTicker.create_perspective(app.FOV, np.float32(W_) / np.float32(H_)),
             app.CameraTicker.create_camera(a, da))
    assert np.array_equal(r.hits(), hits_tick)
    assert np.array_equal(r.radiance().view(np.uint32), rad_tick.view(np.uint32))
    assert (hits_tick["inst"] != 0xFFFFFFFF).mean() > 0.02


def test_main_loop_streams_terrain_one_texture_per_tick():
    r = R.reset_renderer()
    q = R.TextureUploadQueue()
    w = app.WorldState(q, _entities())
    w.camera_position = np.array([0.0, 2.0, -10.0], np.float32)
    r.set_params(scenes.RenderParams(64, 48, mode=0))
    counts = []

    def script(k, ptr):
        ptr.contents.keys[0] = 1                 # W held
        r.lib.vt_set_cursor(3.0 * k, 0.0)        # turning right a little every frame
        counts.append(r.lib.vt_instance_count())

    _say("case: 8 frames of the main loop")
    ticker, given = app.run(r, w, q, frames=8, script=script)
    assert ticker.frame_num == 8
    # textures arrive one per tick (render.rs:246): AncientTemple (60 of the 11 x 11 entities),
    # then Treasure (the other 61), then one terrain chunk per frame; none before (0 -> 1 record)
    assert counts == [1, 60, 121] + [121 + (k - 2) for k in range(3, 8)], counts
    thetas = [float(np.arctan2(d[2], d[0])) for _, d in given]
    assert thetas[-1] > thetas[2]
