"""Headless host loop on the CPU: world.rs camera movement and scene, render.rs camera latency,
main.rs ordering (with a stand-in renderer object; the GPU path is in test_gpu_app.py)."""
import ctypes as C

import numpy as np

from vtrace_amd import app
from vtrace_amd import renderer as R
from vtrace_amd import world as W
from vtrace_amd.abi import UserInput

F32 = np.float32


def _entities():
    a = np.zeros((8, 8, 8), np.uint32)
    a[2:6, 2:6, 2:6] = 0xFF00FF00
    b = np.zeros((8, 8, 8), np.uint32)
    b[:, :4, :] = 0xFF0000FF
    from vtrace_amd import scenes
    return scenes.Volume(a, None), scenes.Volume(b, None)


def _world():
    q = R.TextureUploadQueue()
    return app.WorldState(q, _entities()), q


def test_initial_state_and_direction():
    w, q = _world()
    assert [h for _, h in q.queue] == [0, 1]          # AncientTemple, Treasure (world.rs:45-46)
    assert w.camera_phi == F32(np.float32(np.pi) / F32(2.0))
    d = w.get_camera_direction()
    assert d[0] == F32(1.0) and abs(float(d[1])) < 1e-7 and d[2] == F32(0.0)


def test_keys_move_along_view_frame():
    w, q = _world()
    keys = {"w": (1, 0, 0, 0, 0, 0), "a": (0, 1, 0, 0, 0, 0), "s": (0, 0, 1, 0, 0, 0),
            "d": (0, 0, 0, 1, 0, 0), "space": (0, 0, 0, 0, 1, 0), "lshift": (0, 0, 0, 0, 0, 1)}
    want = {"w": (2.5, 0, 0), "a": (0, 0, -2.5), "s": (-2.5, 0, 0), "d": (0, 0, 2.5),
            "space": (0, -2.5, 0), "lshift": (0, 2.5, 0)}
    for k, keyset in keys.items():
        w.camera_position = np.zeros(3, np.float32)
        w.camera_theta = F32(0.0)
        w.move(0.5, app.Input(keyset))
        assert np.array_equal(w.camera_position, np.array(want[k], np.float32)), k
    # theta = pi/2: W moves along +z
    w.camera_position = np.zeros(3, np.float32)
    w.camera_theta = F32(np.pi / 2)
    w.move(1.0, app.Input(keys["w"]))
    assert abs(float(w.camera_position[2]) - 5.0) < 1e-6 and abs(float(w.camera_position[0])) < 1e-5


def test_mouse_wraps_theta_and_clamps_phi():
    w, _ = _world()
    w.move(0.0, app.Input(mouse=(10.0, 0.0), last_mouse=(0.0, 0.0)))
    assert w.camera_theta == F32(0.0) + F32(0.02) * F32(10.0)
    w.camera_theta = F32(0.0)
    w.move(0.0, app.Input(mouse=(-10.0, 0.0)))
    assert w.camera_theta == F32(F32(-0.2) + F32(2.0) * app.PI)
    w.move(0.0, app.Input(mouse=(0.0, 1000.0)))
    assert w.camera_phi == F32(0.1)
    w.move(0.0, app.Input(mouse=(0.0, -1000.0)))
    assert w.camera_phi == F32(app.PI - F32(0.1))
    w.accum_time_frac = F32(0.75)
    w.move(0.5, app.Input())
    assert w.accum_time_whole == 1 and w.accum_time_frac == F32(F32(1.25) - F32(1.0))


def test_scene_entities_then_terrain():
    w, q = _world()
    scene = w.update(0.0, app.Input(), q)
    kids = list(scene)
    ent = kids[:121]
    assert len(kids) == 121 + (len(q.queue) - 2)
    t_handle, temple_handle = w.entity_texture_registry["Treasure"], w.entity_texture_registry["AncientTemple"]
    k = 0
    for x in range(-5, 6):
        for z in range(-5, 6):
            m, h = ent[k]
            M = m.reshape(4, 4)
            assert np.array_equal(M[3, :3], [F32(x) * F32(1.5), -5.0, F32(z) * F32(1.5)])
            assert h == (t_handle if (x + z + 10) % 2 == 0 else temple_handle)
            k += 1
    # the terrain part is the pager's scene for the same camera
    q2 = R.TextureUploadQueue()
    ref = W.terrain_scene(w.camera_position, W.WorldPager(0), q2)
    assert [np.array_equal(a[0], b[0]) for a, b in zip(kids[121:], ref)] == [True] * len(kids[121:])
    assert w.frame_num == 1


class _StubRenderer:
    """Records the matrices render_tick is given; 640 x 480 window."""

    def __init__(self):
        self.calls = []
        self.window_width = self.window_height = 1
        self.input = UserInput()
        self.instances = []

        class _Lib:
            pass
        self.lib = _Lib()
        self.lib.get_input_data_pointer = lambda: C.pointer(self.input)

    def render_tick(self, P, Cm, queue=None, uploads_per_tick=1):
        self.calls.append((np.array(P), np.array(Cm)))
        if queue is not None and uploads_per_tick == 1:
            queue.pop()
        self.window_width, self.window_height = 640, 480
        return True

    def update_instances(self, scene):
        self.instances.append(scene.num_total_children())


def test_one_frame_camera_latency_and_resize():
    r = _StubRenderer()
    t = app.CameraTicker(r)
    a, b = np.array([1, 2, 3], np.float32), np.array([4, 5, 6], np.float32)
    da = np.array([0, 0, 1], np.float32)
    t.tick(a, da, None)
    t.tick(b, da, None)
    t.tick(b, da, None)
    P0 = app.CameraTicker.create_perspective(app.FOV, 1.0)
    P1 = app.CameraTicker.create_perspective(app.FOV, F32(640) / F32(480))
    C0 = app.CameraTicker.create_camera(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32))
    assert np.array_equal(r.calls[0][0], P0) and np.array_equal(r.calls[0][1], C0)
    assert np.array_equal(r.calls[1][0], P1)
    assert np.array_equal(r.calls[1][1], app.CameraTicker.create_camera(a, da))
    assert np.array_equal(r.calls[2][1], app.CameraTicker.create_camera(b, da))


def test_run_loop_order_and_keys():
    r = _StubRenderer()
    w, q = _world()

    def script(k, ptr):
        ptr.contents.keys[0] = 1 if k < 3 else 0      # hold W for three frames

    ticker, given = app.run(r, w, q, frames=5, script=script)
    assert len(r.calls) == 5 and len(r.instances) == 5
    xs = [float(p[0]) for p, _ in given]
    assert xs[0] == 0.0 and xs[1] == 0.0               # dt of the first update is 0
    assert xs[2] > xs[1] and xs[3] > xs[2] and xs[4] == xs[3]   # W moves +x, released at frame 3
