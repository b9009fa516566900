"""Benchmark: Mrays/s of the plain path on the 1080p 4-spp path-tracing workload.

Workload (one MI355X): 16 noise volumes of 256^3 (fill U[10%,40%]) at non-overlapping random
places in a 2048^3 world, 1920x1080, 4 spp, up to 4 cosine-weighted diffuse bounces plus one
sun shadow ray per vertex, hash-RNG seeded per pixel/sample, FP32 radiance.

A step is one frame: LBVH traversal + brick DDA + shading for every pixel sample, with the
volumes, LBVH and instances already resident in HBM.  Rays counted = traversal queries
(primary + shadow + bounce) from the library's own counters on an untimed counting frame.

Roofline: algorithmic bytes per frame = 32 B per LBVH node visited + 64 B per instance
entered + (4 B + 64 B if non-empty) per brick entered + 5 B per hit + 32 B per pixel,
divided by the trace kernel's mean duration (HIP events on the library stream), against
the MI355X HBM peak of 8 TB/s.

cpu_baseline: the CPU oracle (a scalar FP32/FP64 restatement of the same path, one core)
on a band of rows of the same frame, scaled to Mrays/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(st: dict, pixels: int) -> int:
    return (32 * st["nodes"] + 64 * st["instances"] + 4 * st["bricks"] + 64 * st["bricks_nonempty"] + 5 * st["hits"]
            + 32 * pixels)


def cpu_baseline(scene, target_s: float = 15.0) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O  # the oracle is the CPU baseline's timed code, never the product

    H = scene.params.height
    rows, elapsed, rays = 1, 0.0, 0
    y0 = H // 2
    # grow the band until the sample takes a measurable time, then run the final sample
    while True:
        t0 = time.perf_counter()
        _, _, st = O.render(scene, use_bvh=True, rows=(y0, min(H, y0 + rows)))
        dt = time.perf_counter() - t0
        if dt > target_s / 8 or y0 + rows >= H:
            elapsed, rays = dt, st["rays"]
            break
        rows = min(H - y0, max(rows * 2, int(rows * (target_s / 8) / max(dt, 1e-3))))
    rows_f = min(H, y0 + rows) - y0
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "port",
            "sample": f"oracle (scalar C++, -O2) on rows {y0}..{y0 + rows_f} of the same frame "
                      f"({rays} rays, {elapsed:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the renderer has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    from vtrace_amd import renderer as R
    from vtrace_amd import scenes

    t0 = time.perf_counter()
    sc = scenes.config2(args.width, args.height)
    t_gen = time.perf_counter() - t0
    r = R.get_renderer()
    t0 = time.perf_counter()
    R.load_scene(r, sc, stats=True)
    t_up = time.perf_counter() - t0
    print(f"# scene {sc.name}: generated in {t_gen:.1f} s, uploaded+bricked in {t_up:.1f} s", file=sys.stderr,
          flush=True)
    # counting frame (untimed): rays and algorithmic bytes
    r.render(sc.P, sc.C)
    st = r.stats()
    pixels = sc.params.width * sc.params.height
    bytes_frame = algorithmic_bytes(st, pixels)
    hits = r.hits()
    hit_frac = float((hits["inst"] != 0xFFFFFFFF).mean())
    r.set_params(sc.params, stats=False)
    r.render(sc.P, sc.C)
    # timed: K frames, kernel time from HIP events on the library stream, wall clock around
    r.lib.vt_synchronize()
    t0 = time.perf_counter()
    kernel_ms, wall_ms = r.bench(args.warmup, args.steps)
    r.lib.vt_synchronize()
    t_all = time.perf_counter() - t0
    # per-frame latency distribution (render_tick round trips)
    lat = []
    for _ in range(min(args.steps, 20)):
        r.render(sc.P, sc.C)
        lat.append(r.last_frame_ms())
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]
    mrays = st["rays"] / (wall_ms * 1e-3) / 1e6
    achieved = bytes_frame / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s @1080p 4spp; p50/p99 render latency; time to relMSE 1e-3 vs plain",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded value-noise volumes, seeded camera)",
        "config": {"workload": sc.name, "instances": 16, "volume": "256^3 u8 index + palette", "world": 2048,
                   "width": sc.params.width, "height": sc.params.height, "spp": sc.params.spp,
                   "max_bounces": sc.params.max_bounces, "rays_per_frame": st["rays"],
                   "primary_hit_fraction": round(hit_frac, 4), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_ray": round(bytes_frame / max(1, st["rays"]), 1),
                     "kernel": "k_render<2,false>", "kernel_ms": round(kernel_ms, 4)},
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4)},
        "counters": st,
        "timed_region_s": round(t_all, 3),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
