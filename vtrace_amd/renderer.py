"""Host-side mirror of the reference's renderer client (src/render.rs, src/scene.rs).

``Renderer`` drives librender_hip through the same call sequence ``src/render.rs`` uses on
the Vulkan library: ``entry`` at construction, ``start_update_instances`` /
``end_update_instances`` with one record per uploaded child (un-uploaded textures are
skipped, render.rs:220-238), at most one ``add_texture`` per ``render_tick``
(render.rs:246-268; ``uploads_per_tick=None`` drains the queue with batched uploads instead),
then ``render_tick`` with the perspective and camera matrices.
``SceneGraph`` flattens parents and children depth-first in child order; instance order
is the draw order and therefore the instance index of the hit buffers (scene.rs:146-181).

The headless extensions (render mode, index+palette volumes, read-back) are exposed as
methods.  Nothing here computes pixels: every frame is rendered by the HIP library.
"""
from __future__ import annotations

import ctypes as C
from collections import deque

import numpy as np

from . import scenes
from ._native import render_lib
from .abi import RenderTickInfo, VtRenderParams

HIT_DTYPE = np.dtype([("inst", "<u4"), ("voxel", "<u4"), ("steps", "<u2"), ("face", "u1"), ("pal", "u1"),
                      ("mask", "u1"), ("flags", "u1"), ("vhi", "<u2")])


def hit_voxels(hits: np.ndarray) -> np.ndarray:
    """(..., 3) voxel coordinates of hit records: bits 0-9 from `voxel`, bits 10-14 from `vhi`.
    A coordinate of 32768 or more (flags bit 2) comes back as its low 15 bits."""
    v = hits["voxel"].astype(np.int64)
    hi = hits["vhi"].astype(np.int64)
    return np.stack([(v & 1023) | ((hi & 31) << 10), ((v >> 10) & 1023) | (((hi >> 5) & 31) << 10),
                     ((v >> 20) & 1023) | (((hi >> 10) & 31) << 10)], axis=-1)


INSTANCE_DTYPE = np.dtype([("M", "<f4", 12), ("inv", "<f4", 12), ("tex", "<u4"), ("valid", "<u4"),
                           ("size", "<f4", 3), ("dim", "<u4", 3), ("pad", "<u4", 2)])
NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("left", "<u4"), ("right", "<u4")])
STAT_NAMES = ("rays", "nodes", "instances", "bricks", "bricks_nonempty", "hits", "steps")


class RendererError(RuntimeError):
    pass


class SceneGraph:
    """scene.rs:19-29 Parent/Child tree; iteration = DFS pre-order with accumulated models."""

    def __init__(self, model=None, texture_handle=None, children=None):
        self.model = np.eye(4, dtype=np.float32).reshape(16) if model is None else np.asarray(model, np.float32)
        self.texture_handle = texture_handle
        self.children = children if children is not None else ([] if texture_handle is None else None)

    @classmethod
    def child(cls, model, texture_handle):
        return cls(model, texture_handle, None)

    def add_child(self, child: "SceneGraph"):
        if self.children is None:  # a Child becomes a Parent keeping itself as first child (scene.rs:94-115)
            me = SceneGraph.child(np.eye(4, dtype=np.float32).reshape(16), self.texture_handle)
            self.children = [me]
            self.texture_handle = None
        self.children.append(child)

    def num_total_children(self) -> int:
        if self.children is None:
            return 1
        return sum(c.num_total_children() for c in self.children)

    def __iter__(self):
        # scene.rs:146-181: a parent pushes its accumulated model on a stack that is never
        # popped, so siblings after a subtree see the last pushed model (reproduced).
        stack_nodes = [self]
        model_stack = []
        while stack_nodes:
            node = stack_nodes.pop()
            if node.children is not None:
                stack_nodes.extend(reversed(node.children))
                m = node.model.reshape(4, 4)
                model_stack.append(m if not model_stack else _matmul_cm(model_stack[-1], m))
                continue
            m = node.model.reshape(4, 4)
            yield (m if not model_stack else _matmul_cm(model_stack[-1], m)).reshape(16), node.texture_handle


def _matmul_cm(a, b):
    """Column-major product a*b of (4,4) arrays stored as m[c][r]."""
    return (b.astype(np.float32) @ a.astype(np.float32)).astype(np.float32)


class TextureUploadQueue:
    """render.rs:137-171: handles are issued at once, uploads happen one per tick."""

    def __init__(self):
        self.num_textures_created = 0
        self.queue = deque()

    def add_texture(self, volume: "scenes.Volume") -> int:
        handle = self.num_textures_created
        self.num_textures_created += 1
        self.queue.append((volume, handle))
        return handle

    def pop(self):
        return self.queue.popleft() if self.queue else None


class Renderer:
    """render.rs:173-320 over librender_hip."""

    def __init__(self):
        self.lib = render_lib()
        code = self.lib.entry()
        if code != 0:
            raise RendererError(f"ERROR: initialization failed ({code:#x}): {self.error()}")
        self.texture_handle_lookup: dict[int, int] = {}
        self.window_width = 1
        self.window_height = 1

    def error(self) -> str:
        e = self.lib.vt_last_error()
        return e.decode() if e else ""

    def close(self):
        self.lib.cleanup()

    # -- reference call sequence ------------------------------------------------------
    def update_instances(self, scene: SceneGraph):
        count = scene.num_total_children()
        ptr = self.lib.start_update_instances(count)
        if not ptr:
            raise RendererError("ERROR: Updating instances failed")
        buf = np.ctypeslib.as_array(ptr, shape=(max(count, 1) * 16,))
        models, tex = [], []
        for model, handle in scene:
            t = self.texture_handle_lookup.get(handle)
            if t is not None:
                models.append(model)
                tex.append(t)
        n = len(models)
        if n:
            recs = np.asarray(models, dtype=np.float32).reshape(n, 16).copy()
            recs.view(np.uint32)[:, 15] = np.asarray(tex, dtype=np.uint32)  # GPUInstance::from_model
            buf[:16 * n] = recs.reshape(-1)
        if self.lib.end_update_instances(n) != 0:
            raise RendererError("ERROR: Updating instances failed")

    def render_tick(self, perspective: np.ndarray, camera: np.ndarray, queue: TextureUploadQueue | None = None,
                    uploads_per_tick: int | None = 1) -> bool:
        """uploads_per_tick = 1 is the reference (render.rs:246); None drains the queue, k takes
        up to k, uploading runs of same-size RGBA8 volumes together (vt_add_textures)."""
        if queue is not None:
            items = []
            while uploads_per_tick is None or len(items) < uploads_per_tick:
                item = queue.pop()
                if item is None:
                    break
                items.append(item)
            if len(items) == 1:
                volume, handle = items[0]
                self.texture_handle_lookup[handle] = self.upload(volume)
            elif items:
                ids = self.upload_many([v for v, _ in items])
                for (_, handle), tid in zip(items, ids):
                    self.texture_handle_lookup[handle] = tid
        P = np.ascontiguousarray(perspective, dtype=np.float32)
        Cm = np.ascontiguousarray(camera, dtype=np.float32)
        info = RenderTickInfo(P.ctypes.data, Cm.ctypes.data)
        w, h = C.c_int32(0), C.c_int32(0)
        code = self.lib.render_tick(C.byref(w), C.byref(h), C.byref(info))
        self.window_width, self.window_height = w.value, h.value
        return code == 0

    # -- uploads ----------------------------------------------------------------------------
    def add_texture_rgba(self, rgba: np.ndarray) -> int:
        """add_texture (lib/memory.c:286): RGBA8 volume shaped (d, h, w) of uint32 or (d, h, w, 4) uint8."""
        a = np.ascontiguousarray(rgba)
        if a.dtype == np.uint32:
            d, h, w = a.shape
        else:
            d, h, w, _ = a.shape
        tid = self.lib.add_texture(a.ctypes.data, w, h, d)
        if tid < 0:
            raise RendererError(f"ERROR: Adding texture failed: {self.error()}")
        return tid

    def add_volume_indexed(self, idx: np.ndarray, palette: np.ndarray) -> int:
        a = np.ascontiguousarray(idx, dtype=np.uint8)
        pal = np.ascontiguousarray(palette, dtype=np.uint32)
        d, h, w = a.shape
        tid = self.lib.vt_add_volume_indexed(a.ctypes.data, w, h, d, pal.ctypes.data)
        if tid < 0:
            raise RendererError(f"ERROR: Adding volume failed: {self.error()}")
        return tid

    def upload(self, volume: "scenes.Volume") -> int:
        if volume.fmt == 0:
            return self.add_volume_indexed(volume.data, volume.palette)
        return self.add_texture_rgba(volume.data)

    def add_textures_rgba(self, volumes: list) -> list[int]:
        """vt_add_textures: RGBA8 volumes of one shape (d, h, w) uint32, uploaded together."""
        a = np.ascontiguousarray(np.stack([np.asarray(v, dtype=np.uint32) for v in volumes]))
        n, d, h, w = a.shape
        ids = np.zeros(n, np.int32)
        if self.lib.vt_add_textures(a.ctypes.data, n, w, h, d, ids.ctypes.data) != 0:
            raise RendererError(f"ERROR: Adding textures failed: {self.error()}")
        return [int(i) for i in ids]

    def upload_many(self, volumes: list) -> list[int]:
        """Uploads in order; consecutive RGBA8 volumes of one shape go up together."""
        ids: list[int] = []
        k = 0
        while k < len(volumes):
            v = volumes[k]
            if v.fmt == 0:
                ids.append(self.upload(v))
                k += 1
                continue
            j = k + 1
            while j < len(volumes) and volumes[j].fmt == 1 and volumes[j].data.shape == v.data.shape:
                j += 1
            if j - k == 1:
                ids.append(self.add_texture_rgba(v.data))
            else:
                ids.extend(self.add_textures_rgba([u.data for u in volumes[k:j]]))
            k = j
        return ids

    def lbvh_builds(self) -> int:
        return int(self.lib.vt_lbvh_builds())

    def set_instances(self, records: np.ndarray):
        recs = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 16)
        n = recs.shape[0]
        ptr = self.lib.start_update_instances(n)
        if not ptr:
            raise RendererError("start_update_instances failed")
        buf = np.ctypeslib.as_array(ptr, shape=(max(n, 1) * 16,))
        buf[: n * 16] = recs.reshape(-1)
        if self.lib.end_update_instances(n) != 0:
            raise RendererError(f"end_update_instances failed: {self.error()}")

    def refit_instances(self, records: np.ndarray):
        """Moved instances, same count as the last committed list: vt_refit_instances."""
        recs = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 16)
        n = recs.shape[0]
        ptr = self.lib.start_update_instances(n)
        if not ptr:
            raise RendererError("start_update_instances failed")
        buf = np.ctypeslib.as_array(ptr, shape=(max(n, 1) * 16,))
        buf[: n * 16] = recs.reshape(-1)
        if self.lib.vt_refit_instances(n) != 0:
            raise RendererError(f"vt_refit_instances failed: {self.error()}")

    # -- headless extensions ----------------------------------------------------------------
    def set_params(self, p: "scenes.RenderParams", stats: bool = False):
        s = VtRenderParams()
        s.width, s.height, s.viewport_square = p.width, p.height, p.viewport_square
        s.mode, s.spp, s.max_bounces, s.jitter = p.mode, p.spp, p.max_bounces, p.jitter
        s.seed, s.frame = p.seed & 0xFFFFFFFF, p.frame & 0xFFFFFFFF
        for k in range(3):
            s.sun_dir[k], s.sun_radiance[k], s.sky[k] = p.sun_dir[k], p.sun_radiance[k], p.sky[k]
        s.stats = 1 if stats else 0
        if self.lib.vt_set_params(C.byref(s)) != 0:
            raise RendererError(f"vt_set_params failed: {self.error()}")
        self._wh = (p.width, p.height)

    def render(self, P, Cm) -> None:
        if not self.render_tick(P, Cm):
            raise RendererError(f"render_tick failed: {self.error()}")

    def hits(self) -> np.ndarray:
        w, h = self.window_width, self.window_height
        out = np.zeros(w * h, dtype=HIT_DTYPE)
        if self.lib.vt_read_hits(out.ctypes.data, out.nbytes) != 0:
            raise RendererError(self.error())
        return out.reshape(h, w)

    def radiance(self, out: np.ndarray | None = None) -> np.ndarray:
        """(h, w, 4) float32 radiance of the last frame; into `out` (C-contiguous float32 of that
        shape, e.g. a view of a page-locked buffer) when given."""
        w, h = self.window_width, self.window_height
        if out is None:
            out = np.zeros((h, w, 4), dtype=np.float32)
        elif out.shape != (h, w, 4) or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("radiance: out must be C-contiguous float32 of shape (height, width, 4)")
        if self.lib.vt_read_radiance(out.ctypes.data, out.nbytes) != 0:
            raise RendererError(self.error())
        return out

    def enable_gbuffer(self, on: bool = True) -> None:
        if self.lib.vt_enable_gbuffer(1 if on else 0) != 0:
            raise RendererError(self.error())

    def gbuffer(self) -> np.ndarray:
        """(h, w, 8) float32: albedo rgb, distance, normal xyz, hit flag (sample 0 of each pixel)."""
        w, h = self.window_width, self.window_height
        out = np.zeros((h, w, 8), dtype=np.float32)
        if self.lib.vt_read_gbuffer(out.ctypes.data, out.nbytes) != 0:
            raise RendererError(self.error())
        return out

    def bgra8(self) -> np.ndarray:
        w, h = self.window_width, self.window_height
        out = np.zeros((h, w), dtype=np.uint32)
        if self.lib.vt_read_bgra8(out.ctypes.data, out.nbytes) != 0:
            raise RendererError(self.error())
        return out

    def stats(self) -> dict:
        out = np.zeros(7, dtype=np.uint64)
        if self.lib.vt_read_stats(out.ctypes.data) != 0:
            raise RendererError(self.error())
        return {k: int(v) for k, v in zip(STAT_NAMES, out)}

    STAGE_KINDS = ("render", "primary", "shadow", "bounce", "accumulate", "secondary")

    def stage_stats(self) -> list:
        """Per-launch-stage counters of the last counted frame: [(kind, {counter: value})]."""
        out = np.zeros((64, 7), dtype=np.uint64)
        kinds = np.zeros(64, dtype=np.uint8)
        n = self.lib.vt_read_stage_stats(out.ctypes.data, 64, kinds.ctypes.data)
        if n < 0:
            raise RendererError(self.error())
        return [(self.STAGE_KINDS[int(kinds[i])], {k: int(v) for k, v in zip(STAT_NAMES, out[i])}) for i in range(n)]

    def bench_stages(self, warmup: int, iters: int):
        """(kernel_ms, wall_ms, [(kind, mean ms)]) over `iters` frames."""
        k, w = C.c_float(0), C.c_float(0)
        sm = np.zeros(64, dtype=np.float32)
        kinds = np.zeros(64, dtype=np.uint8)
        n = self.lib.vt_bench_stages(warmup, iters, C.byref(k), C.byref(w), sm.ctypes.data, 64, kinds.ctypes.data)
        if n < 0:
            raise RendererError(self.error())
        return k.value, w.value, [(self.STAGE_KINDS[int(kinds[i])], float(sm[i])) for i in range(n)]

    def last_frame_ms(self) -> float:
        return float(self.lib.vt_last_frame_ms())

    def bench(self, warmup: int, iters: int):
        k, w = C.c_float(0), C.c_float(0)
        if self.lib.vt_bench_frames(warmup, iters, C.byref(k), C.byref(w)) != 0:
            raise RendererError(self.error())
        return k.value, w.value

    def volume_dense(self, tid: int) -> np.ndarray:
        info = np.zeros(8, dtype=np.uint32)
        if self.lib.vt_volume_info(tid, info.ctypes.data) != 0:
            raise RendererError("bad volume id")
        w, h, d, fmt = (int(x) for x in info[:4])
        out = np.zeros((d, h, w), dtype=np.uint8 if fmt == 0 else np.uint32)
        if self.lib.vt_read_volume(tid, out.ctypes.data, out.nbytes) != 0:
            raise RendererError(self.error())
        return out

    def volume_info(self, tid: int) -> dict:
        info = np.zeros(8, dtype=np.uint32)
        if self.lib.vt_volume_info(tid, info.ctypes.data) != 0:
            raise RendererError("bad volume id")
        return {"dims": tuple(int(x) for x in info[:3]), "fmt": int(info[3]), "nbricks": int(info[4]),
                "bytes": int(info[5]) | (int(info[6]) << 32)}

    def brick_map(self, tid: int) -> np.ndarray:
        info = self.volume_info(tid)
        bd = [(d + 7) // 8 for d in info["dims"]]
        out = np.zeros(bd[0] * bd[1] * bd[2], dtype=np.uint32)
        if self.lib.vt_read_brick_map(tid, out.ctypes.data, out.nbytes) != 0:
            raise RendererError("vt_read_brick_map failed")
        return out

    def lbvh(self):
        n = self.lib.vt_instance_count()
        nodes = np.zeros(max(1, 2 * n - 1), dtype=NODE_DTYPE)
        inst = np.zeros(max(1, n), dtype=INSTANCE_DTYPE)
        if self.lib.vt_read_lbvh(nodes.ctypes.data, nodes.nbytes, inst.ctypes.data, inst.nbytes) != 0:
            raise RendererError(self.error())
        return nodes, inst


def load_scene(r: Renderer, scene: "scenes.Scene", stats: bool = False) -> list:
    """Uploads a Scene's volumes and instances, sets its parameters; returns texture ids."""
    ids = [r.upload(v) for v in scene.volumes]
    recs = np.array(scene.records, dtype=np.float32).copy()
    # scene records reference volumes by position; remap to the library's texture ids
    tex = recs.view(np.uint32)[:, 15].copy()
    recs.view(np.uint32)[:, 15] = np.array([ids[t] if t < len(ids) else t for t in tex], dtype=np.uint32)
    r.set_instances(recs)
    r.set_params(scene.params, stats=stats)
    return ids


_singleton: Renderer | None = None


def get_renderer() -> Renderer:
    """The library keeps global state (like the reference's `glbl`); one Renderer per process."""
    global _singleton
    if _singleton is None:
        _singleton = Renderer()
    return _singleton


def reset_renderer() -> Renderer:
    global _singleton
    if _singleton is not None:
        _singleton.close()
        _singleton = None
    return get_renderer()
