"""Streamed terrain world: the pager and terrain scene of src/gen/pager.rs and src/world.rs.

``WorldPager.page`` follows pager.rs:52-78: a chunk position is generated once; an occupied
chunk is queued for upload on a ``TextureUploadQueue`` and its handle cached, an empty one is
remembered as ``None``.  ``terrain_scene`` follows world.rs:163-198: the (2*10+1)^3 chunk
positions around the camera's chunk, x outermost, each occupied chunk an instance with model
``translate(chunk * 2.0) * scale(2.0)``.  Voxels come from libvtscene's terrain generator
(terrain.rs:37-94 with value noise in place of noise-rs).

Uploads: the reference's ``render_tick`` uploads one queued texture per tick
(render.rs:246-268); ``Renderer.render_tick(..., uploads_per_tick=None)`` drains the queue
instead, uploading same-size chunks together through ``vt_add_textures``.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import scenes
from ._native import scene_lib
from .renderer import SceneGraph, TextureUploadQueue

CHUNK_VOXEL_SIZE = 16     # pager.rs:24
CHUNK_WORLD_SIZE = 2.0    # pager.rs:25
CHUNK_LOAD_DIST = 10      # pager.rs:26


def get_chunk_pos(pos) -> tuple[int, int, int]:
    """pager.rs:28-34, in FP32 like the reference's Vec3."""
    p = np.asarray(pos, dtype=np.float32)
    q = p / np.float32(CHUNK_WORLD_SIZE)
    return tuple(int(math.floor(float(v))) for v in q)


class TerrainGenerator:
    """terrain.rs:20-94 (seed 0 in world.rs); chunks are (16, 16, 16) uint32 RGBA8, x fastest."""

    def __init__(self, seed: int = 0):
        self.seed = seed & 0xFFFFFFFF
        self.lib = scene_lib()

    def gen_chunk(self, cx: int, cy: int, cz: int) -> scenes.Volume | None:
        out = np.zeros((CHUNK_VOXEL_SIZE,) * 3, np.uint32)
        rc = self.lib.vt_terrain_chunk(cx, cy, cz, self.seed, out.ctypes.data)
        if rc < 0:
            raise RuntimeError("terrain generation failed")
        return scenes.Volume(out, None) if rc == 1 else None

    def gen_chunks(self, positions) -> list[scenes.Volume | None]:
        """Many chunks at once on host threads; same voxels as gen_chunk per position."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.int32).reshape(-1, 3))
        n = pos.shape[0]
        out = np.zeros((n,) + (CHUNK_VOXEL_SIZE,) * 3, np.uint32)
        filled = np.zeros(n, np.uint8)
        if n and self.lib.vt_terrain_chunks(pos.ctypes.data, n, self.seed, out.ctypes.data, filled.ctypes.data) != 0:
            raise RuntimeError("terrain generation failed")
        return [scenes.Volume(out[i], None) if filled[i] else None for i in range(n)]

    def voxel(self, x: int, y: int, z: int) -> int:
        return int(self.lib.vt_terrain_voxel(x, y, z, self.seed))


class WorldPager:
    """pager.rs:39-79."""

    def __init__(self, seed: int = 0):
        self.chunks: dict[tuple[int, int, int], tuple[scenes.Volume, int] | None] = {}
        self.terrain_generator = TerrainGenerator(seed)
        self.view = None  # (camera chunk, its terrain scene): paging is cached, so it cannot change

    def page(self, cx: int, cy: int, cz: int, queue: TextureUploadQueue) -> int | None:
        key = (cx, cy, cz)
        if key in self.chunks:
            hit = self.chunks[key]
            return None if hit is None else hit[1]
        chunk = self.terrain_generator.gen_chunk(cx, cy, cz)
        return self._insert(key, chunk, queue)

    def page_many(self, positions, queue: TextureUploadQueue) -> list[int | None]:
        """page() over positions in order, generating the unseen ones together."""
        positions = [tuple(int(v) for v in p) for p in positions]
        new = []
        seen = set()
        for p in positions:
            if p not in self.chunks and p not in seen:
                seen.add(p)
                new.append(p)
        fresh = dict(zip(new, self.terrain_generator.gen_chunks(new))) if new else {}
        out = []
        for p in positions:
            if p in self.chunks:
                hit = self.chunks[p]
                out.append(None if hit is None else hit[1])
            else:
                out.append(self._insert(p, fresh[p], queue))
        return out

    def _insert(self, key, chunk, queue):
        if chunk is None:
            self.chunks[key] = None
            return None
        handle = queue.add_texture(chunk)
        self.chunks[key] = (chunk, handle)
        return handle


def chunk_model(cx: int, cy: int, cz: int) -> np.ndarray:
    """translate(identity, chunk * 2.0) then scale(2.0) (world.rs:175-192), column-major."""
    m = np.zeros(16, np.float32)
    s = np.float32(CHUNK_WORLD_SIZE)
    m[0] = m[5] = m[10] = s
    m[12], m[13], m[14] = np.float32(cx) * s, np.float32(cy) * s, np.float32(cz) * s
    m[15] = 1.0
    return m


def load_positions(camera_position) -> list[tuple[int, int, int]]:
    """The chunk positions world.rs:163-172 visits, in its loop order (x, then y, then z)."""
    c = get_chunk_pos(camera_position)
    r = range(-CHUNK_LOAD_DIST, CHUNK_LOAD_DIST + 1)
    return [(c[0] + x, c[1] + y, c[2] + z) for x in r for y in r for z in r]


def terrain_scene(camera_position, pager: WorldPager, queue: TextureUploadQueue) -> SceneGraph:
    """world.rs:163-198: scene_terrain for a camera position (chunks paged on the way).  While the
    camera stays in one chunk the visited positions, and so the scene, are those of the last call."""
    cpos = get_chunk_pos(camera_position)
    if pager.view is not None and pager.view[0] == cpos:
        return pager.view[1]
    scene = SceneGraph()
    positions = load_positions(camera_position)
    handles = pager.page_many(positions, queue)
    for p, h in zip(positions, handles):
        if h is not None:
            scene.add_child(SceneGraph.child(chunk_model(*p), h))
    pager.view = (cpos, scene)
    return scene


def _flight(r, eyes, width, height, uploads_per_tick, seed=0):
    """One camera flight through a fresh pager: per tick world.rs's terrain update, then
    update_instances and render_tick (main.rs order).  Returns per-tick wall times (ms) of the
    three parts and the number of chunks still queued after each tick."""
    q = TextureUploadQueue()
    pager = WorldPager(seed)
    P = scenes.perspective(80.0 / 180.0 * 3.1415926, width / height)
    parts = {"world": [], "instances": [], "render_tick": [], "tick": []}
    pending = []
    for eye in eyes:
        t0 = time.perf_counter()
        scene = terrain_scene(eye, pager, q)
        t1 = time.perf_counter()
        r.update_instances(scene)
        t2 = time.perf_counter()
        C = scenes.look_at(eye, (eye[0], eye[1], eye[2] + 1.0))
        if not r.render_tick(P, C, q, uploads_per_tick=uploads_per_tick):
            raise RuntimeError("render_tick failed")
        t3 = time.perf_counter()
        for k, v in (("world", t1 - t0), ("instances", t2 - t1), ("render_tick", t3 - t2), ("tick", t3 - t0)):
            parts[k].append(v * 1e3)
        pending.append(len(q.queue))
    return parts, pending


def benchmark(r, width: int = 1920, height: int = 1080, log=print) -> dict:
    """Streaming terrain: chunk uploads one add_texture per chunk vs vt_add_textures, a camera
    flight at the reference's pacing (one upload per tick) vs draining the queue each tick, and
    the cost of an unchanged vs a changed instance update.  Host wall-clock, GPU synchronised."""
    gen = TerrainGenerator(0)
    t0 = time.perf_counter()
    vols = [v for v in gen.gen_chunks(load_positions((0.0, 0.0, 0.0))) if v is not None]
    gen_ms = (time.perf_counter() - t0) * 1e3
    datas = [v.data for v in vols]
    n = len(datas)
    r.lib.vt_synchronize()
    t0 = time.perf_counter()
    for d in datas:
        r.add_texture_rgba(d)
    single_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    r.add_textures_rgba(datas)
    batch_ms = (time.perf_counter() - t0) * 1e3
    log(f"streaming: {n} chunks: one per call {single_ms:.1f} ms, batched {batch_ms:.1f} ms")

    r.set_params(scenes.RenderParams(width, height, mode=0))
    # fly along +z through the ball of terrain, 1 world unit per tick
    eyes = [(0.5, 1.5, -30.0 + k) for k in range(61)]
    out = {"chunks": n, "generate_ms": round(gen_ms, 2),
           "upload": {"one_per_call_ms": round(single_ms, 2), "batched_ms": round(batch_ms, 2),
                      "chunks_per_s_one_per_call": round(n / (single_ms / 1e3), 1),
                      "chunks_per_s_batched": round(n / (batch_ms / 1e3), 1)},
           "flight": {"ticks": len(eyes), "frame": f"{width}x{height} mode 0"}}
    for name, upt in (("one_upload_per_tick", 1), ("batched", None)):
        parts, pending = _flight(r, eyes, width, height, upt)
        res = {k + "_ms_p50": round(float(np.percentile(v, 50)), 3) for k, v in parts.items()}
        res["tick_ms_max"] = round(float(np.max(parts["tick"])), 3)
        res["total_ms"] = round(float(np.sum(parts["tick"])), 1)
        res["ticks_with_chunks_pending"] = int(sum(1 for n in pending if n))
        res["max_chunks_pending"] = int(max(pending))
        out["flight"][name] = res
        log(f"streaming flight {name}: tick p50 {res['tick_ms_p50']:.2f} ms (world {res['world_ms_p50']:.2f}, "
            f"instances {res['instances_ms_p50']:.2f}, render_tick {res['render_tick_ms_p50']:.2f}), "
            f"{res['ticks_with_chunks_pending']} ticks with chunks pending (max {res['max_chunks_pending']})")
    # instance updates: an unchanged list keeps the tables and tree; a moved list rebuilds
    q = TextureUploadQueue()
    pager = WorldPager(0)
    scene = terrain_scene((0.0, 0.0, -14.0), pager, q)
    r.render_tick(scenes.perspective(1.4, 1.0), scenes.look_at((0, 0, -14), (0, 0, 0)), q, uploads_per_tick=None)
    reps = 20
    r.update_instances(scene)
    t0 = time.perf_counter()
    for _ in range(reps):
        r.update_instances(scene)
    same_ms = (time.perf_counter() - t0) * 1e3 / reps
    moved = SceneGraph()
    for k, (m, h) in enumerate(scene):
        mm = m.copy()
        mm[12] += np.float32(0.001 * (k % 3))
        moved.add_child(SceneGraph.child(mm, h))
    b0 = r.lbvh_builds()
    t0 = time.perf_counter()
    for k in range(reps):
        r.update_instances(moved if k % 2 == 0 else scene)
    changed_ms = (time.perf_counter() - t0) * 1e3 / reps
    out["update_instances_ms"] = {"instances": scene.num_total_children(), "unchanged": round(same_ms, 4),
                                  "changed": round(changed_ms, 4), "rebuilds": int(r.lbvh_builds() - b0)}
    return out
