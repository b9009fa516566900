"""The interactive host loop of the reference, headless: src/world.rs, src/main.rs and the camera
half of src/render.rs over librender_hip.

``WorldState`` follows world.rs:27-205: camera position/angles moved by ``user_input`` in FP32
(W/S along the horizontal view direction, A/D sideways, space down and left shift up, 5 units
per second; mouse deltas times 0.02 on theta and phi, theta wrapped to [0, 2 pi), phi clamped
to [0.1, pi - 0.1]); each update builds the scene of world.rs:137-203: an 11 x 11 grid of the
two entity textures at (1.5 x, -5, 1.5 z), alternating by (x + z + 10) % 2, then the paged
terrain chunks.

``CameraTicker`` follows render.rs:183-320: ``tick(pos, dir, queue)`` uploads (one queued
texture by default), renders with the perspective and camera matrices held from the previous
tick, then rebuilds the perspective if the window size changed and the camera from
(pos, dir) -- the reference's one-frame camera latency -- and reports FPS once a second.

``run`` follows main.rs:24-55: ``update_instances(scene)``, then ``render_tick`` on a second
thread while the world computes the next scene from the input, joined every frame.  Headless,
the host presses keys through ``get_input_data_pointer`` and moves the cursor with
``vt_set_cursor``; render_tick turns the cursor into mouse/last_mouse (lib/entry.c:157-163).

FP32 sin/cos come from numpy; they may differ from the reference's libm results in the last
place (not pinned).
"""
from __future__ import annotations

import math
import threading
import time

import numpy as np

from . import scenes, world
from .renderer import Renderer, SceneGraph, TextureUploadQueue

MOVE_SPEED = np.float32(5.0)     # world.rs:24
SENSITIVITY = np.float32(0.02)   # world.rs:25
PI = np.float32(math.pi)         # world.rs:26
FOV = np.float32(80.0) / np.float32(180.0) * np.float32(3.1415926)   # render.rs:190
F32 = np.float32


class Input:
    """src/render.rs:37-51 by value (what ``unsafe { *input_ptr }`` copies)."""

    __slots__ = ("key_w", "key_a", "key_s", "key_d", "key_space", "key_lshift",
                 "mouse_x", "mouse_y", "last_mouse_x", "last_mouse_y")

    def __init__(self, keys=(0, 0, 0, 0, 0, 0), mouse=(0.0, 0.0), last_mouse=(0.0, 0.0)):
        (self.key_w, self.key_a, self.key_s, self.key_d, self.key_space, self.key_lshift) = keys
        self.mouse_x, self.mouse_y = mouse
        self.last_mouse_x, self.last_mouse_y = last_mouse

    @classmethod
    def from_ptr(cls, p) -> "Input":
        u = p.contents
        return cls(tuple(u.keys), (u.mouse_x, u.mouse_y), (u.last_mouse_x, u.last_mouse_y))


def _v(x, y, z):
    return np.array([x, y, z], dtype=np.float32)


class WorldState:
    """world.rs:27-205.  entity_volumes = the (AncientTemple, Treasure) volumes the reference
    loads from assets/*.vox (world.rs:45-46); they are queued first, as there."""

    def __init__(self, queue: TextureUploadQueue, entity_volumes, seed: int = 0):
        self.camera_position = _v(0.0, 0.0, 0.0)
        self.camera_theta = F32(0.0)
        self.camera_phi = PI / F32(2.0)
        self.accum_time_frac = F32(0.0)
        self.accum_time_whole = 0
        self.frame_num = 0
        temple, treasure = entity_volumes
        self.entity_texture_registry = {"AncientTemple": queue.add_texture(temple),
                                        "Treasure": queue.add_texture(treasure)}
        self.world_pager = world.WorldPager(seed)

    def get_camera_direction(self) -> np.ndarray:
        t, p = self.camera_theta, self.camera_phi
        return _v(np.cos(t) * np.sin(p), np.cos(p), np.sin(t) * np.sin(p))

    def get_horizontal_camera_direction(self) -> np.ndarray:
        return _v(np.cos(self.camera_theta), 0.0, np.sin(self.camera_theta))

    def move(self, dt: float, inp: Input) -> None:
        """world.rs:93-135."""
        dt = F32(dt)
        self.accum_time_frac = F32(self.accum_time_frac + dt)
        if self.accum_time_frac > F32(1.0):
            self.accum_time_frac = F32(self.accum_time_frac - F32(1.0))
            self.accum_time_whole += 1
        t = self.camera_theta
        fwd = _v(np.cos(t), 0.0, np.sin(t))
        side = _v(np.sin(t), 0.0, -np.cos(t))
        up = _v(0.0, 1.0, 0.0)
        pos = self.camera_position
        if inp.key_w > 0:
            pos = pos + fwd * dt * MOVE_SPEED
        if inp.key_s > 0:
            pos = pos - fwd * dt * MOVE_SPEED
        if inp.key_a > 0:
            pos = pos + side * dt * MOVE_SPEED
        if inp.key_d > 0:
            pos = pos - side * dt * MOVE_SPEED
        if inp.key_space > 0:
            pos = pos - up * dt * MOVE_SPEED
        if inp.key_lshift > 0:
            pos = pos + up * dt * MOVE_SPEED
        self.camera_position = pos.astype(np.float32)
        self.camera_theta = F32(self.camera_theta + SENSITIVITY * F32(inp.mouse_x - inp.last_mouse_x))
        self.camera_phi = F32(self.camera_phi - SENSITIVITY * F32(inp.mouse_y - inp.last_mouse_y))
        if self.camera_theta < F32(0.0):
            self.camera_theta = F32(self.camera_theta + F32(2.0) * PI)
        if self.camera_theta >= F32(2.0) * PI:
            self.camera_theta = F32(self.camera_theta - F32(2.0) * PI)
        if self.camera_phi < F32(0.1):
            self.camera_phi = F32(0.1)
        if self.camera_phi >= PI - F32(0.1):
            self.camera_phi = F32(PI - F32(0.1))

    def update(self, dt: float, inp: Input, queue: TextureUploadQueue) -> SceneGraph:
        """world.rs:87-205: move the camera, then the scene (entities, then terrain)."""
        self.move(dt, inp)
        scene = SceneGraph()
        entities = SceneGraph()
        h1 = self.entity_texture_registry["Treasure"]
        h2 = self.entity_texture_registry["AncientTemple"]
        for x in range(-5, 6):
            for z in range(-5, 6):
                m = np.eye(4, dtype=np.float32)
                m[3, 0], m[3, 1], m[3, 2] = F32(x) * F32(1.5), F32(-5.0), F32(z) * F32(1.5)
                entities.add_child(SceneGraph.child(m.reshape(16), h1 if (x + z + 10) % 2 == 0 else h2))
        terrain = world.terrain_scene(self.camera_position, self.world_pager, queue)
        scene.add_child(entities)
        scene.add_child(terrain)
        self.frame_num += 1
        return scene


class CameraTicker:
    """The camera state of render.rs:173-320 around ``Renderer.render_tick``."""

    def __init__(self, r: Renderer, log=None):
        self.r = r
        self.fov = FOV
        self.window_width = self.window_height = 1
        self.prev_window_width = self.prev_window_height = 1
        self.frame_num = 0
        self.perspective = self.create_perspective(self.fov, 1.0)
        self.camera = self.create_camera(_v(0.0, 0.0, 0.0), _v(1.0, 0.0, 0.0))
        self.prev_time = time.perf_counter()
        self.prev_frame_time = time.perf_counter()
        self.prev_frame_num = 0
        self.log = log

    @staticmethod
    def create_perspective(fov, aspect) -> np.ndarray:
        return scenes.perspective(float(fov), float(aspect), 0.01, 10000.0)

    @staticmethod
    def create_camera(position, direction) -> np.ndarray:
        p = np.asarray(position, np.float32)
        return scenes.look_at(p, p + np.asarray(direction, np.float32), (0.0, 1.0, 0.0))

    def tick(self, pos, direction, queue: TextureUploadQueue | None, uploads_per_tick: int | None = 1):
        """render.rs:240-320: returns (ok, seconds since the previous tick)."""
        ok = self.r.render_tick(self.perspective, self.camera, queue, uploads_per_tick=uploads_per_tick)
        self.window_width, self.window_height = self.r.window_width, self.r.window_height
        if self.window_width != self.prev_window_width or self.window_height != self.prev_window_height:
            self.perspective = self.create_perspective(self.fov, F32(self.window_width) / F32(self.window_height))
            self.prev_window_width, self.prev_window_height = self.window_width, self.window_height
        self.camera = self.create_camera(pos, direction)
        now = time.perf_counter()
        dt_us = (now - self.prev_time) * 1e6
        if dt_us > 1e6:
            frames = self.frame_num - self.prev_frame_num
            if self.log is not None and frames:
                self.log(f"FPS: {1e6 * frames / dt_us}   MS: {dt_us / 1000.0 / frames}")
            self.prev_time = now
            self.prev_frame_num = self.frame_num
        frame_dt = now - self.prev_frame_time
        self.prev_frame_time = now
        self.frame_num += 1
        return ok, frame_dt


def run(r: Renderer, world_state: WorldState, queue: TextureUploadQueue, frames: int, script=None,
        uploads_per_tick: int | None = 1, ticker: CameraTicker | None = None):
    """main.rs:24-55 for at most `frames` frames.  script(k, input_ptr) may press keys and move
    the cursor (vt_set_cursor) before frame k.  Returns the CameraTicker and the per-frame
    (camera position, direction) that each render_tick was given."""
    ticker = ticker or CameraTicker(r)
    ptr = r.lib.get_input_data_pointer()
    scene = world_state.update(0.0, Input.from_ptr(ptr), queue)
    code, dt = True, 0.0
    given = []
    for k in range(frames):
        if not code:
            break
        pos = world_state.camera_position.copy()
        direction = world_state.get_camera_direction()
        given.append((pos, direction))
        r.update_instances(scene)
        if script is not None:
            script(k, ptr)
        # main.rs copies *input_ptr while render_tick runs on the other thread (a race); the
        # copy is taken before the thread starts, so a headless run is deterministic
        inp = Input.from_ptr(ptr)
        out = []
        th = threading.Thread(target=lambda: out.append(ticker.tick(pos, direction, queue, uploads_per_tick)))
        th.start()
        scene = world_state.update(dt, inp, queue)
        th.join()
        code, dt = out[0] if out else (False, 0.0)
    return ticker, given
