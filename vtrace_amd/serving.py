"""Serving loop driver: resident scenes, a Poisson request stream, continuous batching.

The scheduler itself is native (vt_serve_* in librender_hip, serve.hip): FIFO admission by
slots and HBM budget, one batched launch per iteration holding sample chunks of the admitted
requests chosen by a policy (round robin, head + riders, shortest remaining work first),
optional hipGraph replay, two iterations in flight.  This module feeds it requests at their
arrival times, measures per-request latency (arrival -> result in host memory) on the wall
clock, and compares against the plain path serving the same stream one request at a time.

The request stream (seeded): resolution uniform over {256x256, 640x480, 1280x720,
1920x1080}, spp uniform in 1..16, scene uniform over the resident scenes, camera pose
seeded per request, Poisson arrivals at a rate set for a target utilisation of the
measured batched capacity.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import scenes
from .renderer import HIT_DTYPE, RendererError

_HOST_BUF = None


def result_buffer(width: int, height: int) -> np.ndarray:
    """A (height, width, 4) float32 view of one reused page-locked host buffer (grown as needed):
    results copied into it go by DMA straight from the device, without the staging copy and the
    first-touch page faults of a fresh pageable array.  The next call's results overwrite it."""
    global _HOST_BUF
    nbytes = width * height * 16
    if _HOST_BUF is None or _HOST_BUF.nbytes < nbytes:
        import torch   # page-locked host memory (torch's pinned allocator)
        _HOST_BUF = torch.empty(max(nbytes, 1920 * 1080 * 16), dtype=torch.uint8, pin_memory=True).numpy()
    return _HOST_BUF[:nbytes].view(np.float32).reshape(height, width, 4)

RESOLUTIONS = ((256, 256), (640, 480), (1280, 720), (1920, 1080))


ROUND_ROBIN, HEAD_RIDERS, SHORTEST_FIRST = 0, 1, 2
POLICY_NAMES = {ROUND_ROBIN: "round_robin", HEAD_RIDERS: "head_riders", SHORTEST_FIRST: "shortest_first"}


class VtServeConfig(C.Structure):
    _fields_ = [("max_slots", C.c_uint32), ("spp_chunk", C.c_uint32), ("memory_budget", C.c_uint64),
                ("max_work", C.c_uint32), ("tile_w", C.c_uint32), ("max_bounces", C.c_uint32),
                ("use_graph", C.c_uint32), ("policy", C.c_uint32), ("pipeline", C.c_uint32), ("streams", C.c_uint32),
                ("poll", C.c_uint32)]


class VtRequest(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("seed", C.c_uint32), ("frame", C.c_uint32),
                ("viewport_square", C.c_uint32), ("perspective", C.c_float * 16), ("camera", C.c_float * 16),
                ("sun_dir", C.c_float * 3), ("sun_radiance", C.c_float * 3), ("sky", C.c_float * 3)]


class VtServeStats(C.Structure):
    _fields_ = [("queued", C.c_uint32), ("active", C.c_uint32), ("done_unread", C.c_uint32), ("tile_w", C.c_uint32),
                ("memory_in_use", C.c_uint64), ("iterations", C.c_uint64), ("samples", C.c_uint64),
                ("last_iteration_ms", C.c_float), ("sub_batches", C.c_uint32)]


def declare(lib):
    if getattr(lib, "_vt_serve_declared", False):
        return
    lib.vt_serve_init.restype = C.c_int32
    lib.vt_serve_init.argtypes = [C.POINTER(VtServeConfig)]
    lib.vt_serve_add_scene.restype = C.c_int32
    lib.vt_serve_add_scene.argtypes = [C.c_void_p, C.c_uint32]
    lib.vt_serve_submit.restype = C.c_int64
    lib.vt_serve_submit.argtypes = [C.POINTER(VtRequest)]
    lib.vt_serve_step.restype = C.c_int32
    lib.vt_serve_step.argtypes = [C.c_void_p, C.c_uint32]
    lib.vt_serve_read.restype = C.c_int32
    lib.vt_serve_read.argtypes = [C.c_int64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
    lib.vt_serve_status.restype = C.c_int32
    lib.vt_serve_status.argtypes = [C.POINTER(VtServeStats)]
    lib.vt_serve_shutdown.restype = None
    lib.vt_serve_shutdown.argtypes = []
    lib._vt_serve_declared = True


@dataclass
class Request:
    scene: int
    width: int
    height: int
    spp: int
    P: np.ndarray
    C: np.ndarray
    seed: int
    max_bounces: int = 4
    sun_dir: tuple = (0.3, -0.8, 0.5)
    sun_radiance: tuple = (2.5, 2.5, 2.5)
    sky: tuple = scenes.SKY
    viewport_square: int = 0
    arrival: float = 0.0

    def params(self) -> "scenes.RenderParams":
        return scenes.RenderParams(self.width, self.height, viewport_square=self.viewport_square, mode=2,
                                   spp=self.spp, max_bounces=self.max_bounces, jitter=1, seed=self.seed,
                                   sun_dir=self.sun_dir, sun_radiance=self.sun_radiance, sky=self.sky)

    def to_c(self) -> VtRequest:
        q = VtRequest()
        q.scene, q.width, q.height, q.spp = self.scene, self.width, self.height, self.spp
        q.max_bounces, q.seed, q.frame, q.viewport_square = self.max_bounces, self.seed & 0xFFFFFFFF, 0, \
            self.viewport_square
        for k in range(16):
            q.perspective[k] = float(self.P[k])
            q.camera[k] = float(self.C[k])
        for k in range(3):
            q.sun_dir[k], q.sun_radiance[k], q.sky[k] = self.sun_dir[k], self.sun_radiance[k], self.sky[k]
        return q


class ServingLoop:
    def __init__(self, renderer, max_slots=32, spp_chunk=4, memory_budget=2 << 30, max_work=16 << 20, tile_w=8,
                 max_bounces=4, use_graph=True, policy=HEAD_RIDERS, pipeline=2, streams=4, poll=False):
        self.r = renderer
        self.lib = renderer.lib
        declare(self.lib)
        cfg = VtServeConfig(max_slots, spp_chunk, memory_budget, max_work, tile_w, max_bounces, 1 if use_graph else 0,
                            policy, pipeline, streams, 1 if poll else 0)
        if self.lib.vt_serve_init(C.byref(cfg)) != 0:
            raise RendererError(f"vt_serve_init failed: {renderer.error()}")
        self.cfg = cfg
        self._done = (C.c_int64 * 64)()

    def add_scene(self, records: np.ndarray) -> int:
        recs = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 16)
        sid = self.lib.vt_serve_add_scene(recs.ctypes.data, recs.shape[0])
        if sid < 0:
            raise RendererError(f"vt_serve_add_scene failed: {self.r.error()}")
        return sid

    def submit(self, req: Request) -> int:
        q = req.to_c()
        rid = self.lib.vt_serve_submit(C.byref(q))
        if rid < 0:
            raise RendererError(f"vt_serve_submit failed: {self.r.error()}")
        return int(rid)

    def step(self) -> list:
        n = self.lib.vt_serve_step(self._done, 64)
        if n < 0:
            raise RendererError(f"vt_serve_step failed: {self.r.error()}")
        return [int(self._done[i]) for i in range(min(n, 64))]

    def read(self, rid: int, width: int, height: int, hits: bool = False, out: np.ndarray | None = None):
        """Radiance (and hit records) of a finished request; radiance into `out` when given
        (C-contiguous float32, (height, width, 4))."""
        rad = np.zeros((height, width, 4), dtype=np.float32) if out is None else out
        if rad.shape != (height, width, 4) or rad.dtype != np.float32 or not rad.flags.c_contiguous:
            raise ValueError("read: out must be C-contiguous float32 of shape (height, width, 4)")
        hb = np.zeros((height, width), dtype=HIT_DTYPE) if hits else None
        rc = self.lib.vt_serve_read(rid, rad.ctypes.data, rad.nbytes, hb.ctypes.data if hits else None,
                                    hb.nbytes if hits else 0)
        if rc != 0:
            raise RendererError(f"vt_serve_read failed for {rid}")
        return (rad, hb) if hits else rad

    def status(self) -> dict:
        st = VtServeStats()
        self.lib.vt_serve_status(C.byref(st))
        return {n: getattr(st, n) for n, _ in VtServeStats._fields_ if n != "pad"}

    def run_all(self, reqs: list, keep: bool = True) -> dict:
        """Submits every request at once and steps until all are done; returns id -> radiance
        (keep = False: each result is read into the reused page-locked buffer and dropped)."""
        ids = {self.submit(q): q for q in reqs}
        out = {}
        pending = set(ids)
        while pending:
            for rid in self.step():
                q = ids[rid]
                if keep:
                    out[rid] = self.read(rid, q.width, q.height)
                else:
                    self.read(rid, q.width, q.height, out=result_buffer(q.width, q.height))
                pending.discard(rid)
        return {list(ids).index(k): v for k, v in out.items()}

    def run_stream(self, reqs: list, timeout_s: float = 600.0) -> dict:
        """Feeds requests at their arrival times (seconds from start); latency = result ready - arrival."""
        order = sorted(range(len(reqs)), key=lambda i: reqs[i].arrival)
        t0 = time.perf_counter()
        nxt = 0
        idmap = {}
        lat = np.zeros(len(reqs))
        done = 0
        samples = 0
        while done < len(reqs):
            now = time.perf_counter() - t0
            if now > timeout_s:
                raise RuntimeError("serving stream timed out")
            while nxt < len(order) and reqs[order[nxt]].arrival <= now:
                i = order[nxt]
                idmap[self.submit(reqs[i])] = i
                nxt += 1
            st = self.status()
            if st["queued"] == 0 and st["active"] == 0:
                if nxt < len(order):  # idle until the next arrival
                    time.sleep(max(0.0, reqs[order[nxt]].arrival - (time.perf_counter() - t0)))
                continue
            for rid in self.step():
                i = idmap[rid]
                q = reqs[i]
                self.read(rid, q.width, q.height, out=result_buffer(q.width, q.height))
                lat[i] = (time.perf_counter() - t0) - q.arrival
                samples += q.width * q.height * q.spp
                done += 1
        wall = time.perf_counter() - t0
        return {"latency_s": lat, "wall_s": wall, "samples": samples}

    def close(self):
        self.lib.vt_serve_shutdown()


def plain_one_at_a_time(renderer, scene_list, reqs: list, timeout_s: float = 600.0) -> dict:
    """The same stream served FIFO by the plain single-frame path, one request at a time."""
    order = sorted(range(len(reqs)), key=lambda i: reqs[i].arrival)
    t0 = time.perf_counter()
    lat = np.zeros(len(reqs))
    cur_scene = None
    for i in order:
        q = reqs[i]
        now = time.perf_counter() - t0
        if now > timeout_s:
            raise RuntimeError("plain stream timed out")
        if q.arrival > now:
            time.sleep(q.arrival - now)
        if cur_scene != q.scene:
            renderer.set_instances(scene_list[q.scene])
            cur_scene = q.scene
        renderer.set_params(q.params())
        renderer.render(q.P, q.C)
        renderer.radiance(out=result_buffer(q.width, q.height))
        lat[i] = (time.perf_counter() - t0) - q.arrival
    return {"latency_s": lat, "wall_s": time.perf_counter() - t0}


# ---- workload -------------------------------------------------------------------------------
def resident_volumes(count=32, sizes=(64, 128, 256, 512), seed=4, fill=(0.05, 0.40)):
    """count seeded noise volumes with sizes drawn from `sizes` and fill U[fill]."""
    rng = scenes.Rng(seed)
    vols = []
    for i in range(count):
        n = sizes[rng.randint(0, len(sizes) - 1)]
        f = rng.uniform(*fill)
        vols.append(scenes.Volume(scenes.gen_volume((n, n, n), seed * 100 + i, f), scenes.gen_palette(seed * 100 + i)))
    return vols


def scene_records(tex_ids, dims) -> list:
    """One instance per scene: volume k at the origin, one world unit per voxel."""
    return [scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), float(n)), t)[None, :]
            for t, n in zip(tex_ids, dims)]


def request_stream(nreq: int, nscenes: int, dims: list, rate_hz: float, seed: int = 5, max_bounces: int = 4,
                   resolutions=RESOLUTIONS, spp_range=(1, 16)) -> list:
    rng = scenes.Rng(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += -math.log(max(1e-12, 1.0 - rng.uniform())) / rate_hz
        w, h = resolutions[rng.randint(0, len(resolutions) - 1)]
        spp = rng.randint(spp_range[0], spp_range[1])
        sc = rng.randint(0, nscenes - 1)
        n = dims[sc]
        eye, target = scenes.orbit_camera(rng, (0.0, 0.0, 0.0), rng.uniform(1.2, 2.0) * n, 0.1 * n)
        out.append(Request(sc, w, h, spp, scenes.perspective(scenes.FOV, w / h), scenes.look_at(eye, target),
                           seed=1000 + i, max_bounces=max_bounces, sun_dir=scenes.sun_direction(rng), arrival=t))
    return out


def percentiles(lat) -> dict:
    lat = np.asarray(lat)
    return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3), "max_ms": round(float(lat.max()) * 1e3, 3)}


# ---- autotuner ------------------------------------------------------------------------------
TILE_CANDIDATES = (4, 8, 16, 32, 64)


def autotune_tile(renderer, probe_reqs: list, scene_recs: list, cache_path: str | None = None, key: str = "",
                  iters: int = 2, **cfg) -> dict:
    """Times the batched loop on `probe_reqs` for every tile shape; returns the fastest.

    Results are cached in cache_path (JSON) under `key` (e.g. the device name and workload)."""
    if cache_path and os.path.exists(cache_path):
        cache = json.load(open(cache_path))
        if key in cache:
            return cache[key]
    timings = {}
    for tw in TILE_CANDIDATES:
        loop = ServingLoop(renderer, tile_w=tw, **cfg)
        for recs in scene_recs:
            loop.add_scene(recs)
        best = float("inf")
        for _ in range(iters):
            t0 = time.perf_counter()
            loop.run_all(probe_reqs)
            best = min(best, time.perf_counter() - t0)
        timings[tw] = best
        loop.close()
    tw = min(timings, key=timings.get)
    res = {"tile_w": tw, "tile": f"{tw}x{64 // tw}", "seconds": {str(k): round(v, 5) for k, v in timings.items()}}
    if cache_path:
        try:
            os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
            cache = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
            cache[key] = res
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1)
        except (OSError, ValueError):
            pass  # an unwritable cache only costs a re-tune next time
    return res


# ---- benchmark ------------------------------------------------------------------------------
def default_tune_cache() -> str:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "vtrace_amd", "autotune.json")


def benchmark(renderer, nreq: int = 2000, util: float = 0.75, seed: int = 11, probe: int = 48, max_slots: int = 32,
              spp_chunk: int = 4, memory_budget: int = 2 << 30, max_work: int = 16 << 20, max_bounces: int = 4,
              use_graph: bool = True, policy: int = HEAD_RIDERS, pipeline: int = 2, streams: int = 4,
              poll: bool = True, over: float = 1.2, tune_cache: str | None = None,
              plain: bool = True, log=None) -> dict:
    """The serving workload: 32 resident noise volumes (64^3..512^3, fill 5-40%) and seeded Poisson
    streams of nreq requests.  Capacities (requests/s with a probe of requests all queued at
    once) are measured for the batching loop and for the plain path; each stream is then served
    by both, at `util` of the loop's capacity, at `util` of the plain path's capacity and at
    `over` times the plain path's capacity (more than it can sustain).
    Latency = arrival -> radiance in host memory (both paths read into one reused page-locked
    buffer), wall clock."""
    say = log or (lambda *a: None)
    t0 = time.perf_counter()
    vols = resident_volumes()
    dims = [v.dims[0] for v in vols]
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    ids = [renderer.upload(v) for v in vols]
    renderer.lib.vt_synchronize()
    t_up = time.perf_counter() - t0
    resident = sum(renderer.volume_info(i)["bytes"] for i in ids)
    del vols
    recs = scene_records(ids, dims)
    say(f"serving: 32 volumes generated in {t_gen:.1f} s, bricked in {t_up:.1f} s, {resident / 2**20:.0f} MiB resident")
    cfg = dict(max_slots=max_slots, spp_chunk=spp_chunk, memory_budget=memory_budget, max_work=max_work,
               max_bounces=max_bounces, policy=policy, pipeline=pipeline, streams=streams, poll=poll)
    probe_reqs = request_stream(probe, len(recs), dims, rate_hz=1.0, seed=seed + 1, max_bounces=max_bounces)
    tune = autotune_tile(renderer, probe_reqs[:12], recs, cache_path=tune_cache or default_tune_cache(),
                         key=f"serve32:{max_bounces}b:{POLICY_NAMES[policy]}:{max_work}:s{streams}", **cfg)
    say(f"serving: tile {tune['tile']} ({tune['seconds']})")
    loop = ServingLoop(renderer, tile_w=tune["tile_w"], use_graph=use_graph, **cfg)
    for rc in recs:
        loop.add_scene(rc)
    loop.run_all(probe_reqs, keep=False)  # warm-up: graph captures of the probe's batch shapes
    t0 = time.perf_counter()
    loop.run_all(probe_reqs, keep=False)
    cap_loop = len(probe_reqs) / (time.perf_counter() - t0)
    at_once = [Request(**{**q.__dict__, "arrival": 0.0}) for q in probe_reqs]
    plain_one_at_a_time(renderer, recs, at_once[:4])
    t0 = time.perf_counter()
    plain_one_at_a_time(renderer, recs, at_once)
    cap_plain = len(probe_reqs) / (time.perf_counter() - t0)
    say(f"serving: capacity {cap_loop:.1f} req/s batched, {cap_plain:.1f} req/s plain (probe of {probe})")
    out = {"requests": nreq, "target_utilisation": util, "capacity_req_s": {"batched": round(cap_loop, 2),
                                                                           "plain": round(cap_plain, 2)},
           "tile": tune["tile"], "resident_volume_mib": round(resident / 2**20, 1),
           "config": {"max_slots": max_slots, "memory_budget_gib": memory_budget / 2**30, "max_work": max_work,
                      "graph": bool(use_graph), "max_bounces": max_bounces, "policy": POLICY_NAMES[policy],
                      "pipeline": pipeline, "sub_batch_streams": streams, "poll": bool(poll)}, "streams": {}}
    for tag, rate in (("at_batched_capacity", util * cap_loop), ("at_plain_capacity", util * cap_plain),
                      ("above_plain_capacity", over * cap_plain)):
        reqs = request_stream(nreq, len(recs), dims, rate_hz=rate, seed=seed, max_bounces=max_bounces)
        st0 = loop.status()
        res = loop.run_stream(reqs)
        st1 = loop.status()
        p = percentiles(res["latency_s"])
        iters = max(1, st1["iterations"] - st0["iterations"])
        e = {"arrival_rate_hz": round(rate, 3),
             "batched": dict(p, wall_s=round(res["wall_s"], 3), throughput_req_s=round(nreq / res["wall_s"], 3),
                             msamples_s=round(res["samples"] / res["wall_s"] / 1e6, 2), iterations=iters,
                             mean_batch_samples=int((st1["samples"] - st0["samples"]) / iters))}
        say(f"serving [{tag}, {rate:.1f} req/s]: batched p50 {p['p50_ms']} ms p99 {p['p99_ms']} ms")
        if plain:
            pr = plain_one_at_a_time(renderer, recs, reqs)
            pp = percentiles(pr["latency_s"])
            e["plain"] = dict(pp, wall_s=round(pr["wall_s"], 3), throughput_req_s=round(nreq / pr["wall_s"], 3),
                              msamples_s=round(res["samples"] / pr["wall_s"] / 1e6, 2))
            say(f"serving [{tag}]: plain p50 {pp['p50_ms']} ms p99 {pp['p99_ms']} ms")
        out["streams"][tag] = e
    loop.close()
    return out
