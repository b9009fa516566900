"""Benchmark: Mrays/s of the plain path on the 1080p 4-spp path-tracing workload.

Workload (one MI355X): 16 noise volumes of 256^3 (fill U[10%,40%]) at non-overlapping random
places in a 2048^3 world, 1920x1080, 4 spp, up to 4 cosine-weighted diffuse bounces plus one
sun shadow ray per vertex, hash-RNG seeded per pixel/sample, FP32 radiance.

A step is one frame: LBVH traversal + brick DDA + shading for every pixel sample, with the
volumes, LBVH and instances already resident in HBM.  Rays counted = traversal queries
(primary + shadow + bounce) from the library's own counters on an untimed counting frame.
`value` = rays per frame / mean host time per frame over the K timed frames, rendered as
render_tick renders them (three chunks of tile rows on three streams, the library's default,
each frame finished before the next).  The launch-stage table and the roofline come from the
same frames on one stream with HIP events around every launch (`single_stream_ms_per_frame`).

Roofline (HBM-bound gather work, no MFMA): algorithmic bytes = 32 B per LBVH node visited
+ 64 B per instance entered + (4 B + 64 B if non-empty) per 8^3 brick entered + 5 B per hit,
plus 16 B per pixel for the hit record (primary stage) and 16 B per pixel of radiance
(accumulate stage).  They are counted per launch stage; the dominant kernel is the stage
kind with the largest share of frame time, and its achieved GB/s = its algorithmic bytes
per launch / its mean launch duration (HIP events around each launch).  `traffic` is the
HBM bytes per launch of that kernel from rocprofv3 FETCH_SIZE + WRITE_SIZE, read from
profiles/traffic.json when present (null otherwise).

cpu_baseline: the CPU oracle (C++ restatement of the same path, host threads over rows) on a
band of rows of the same frame, scaled to Mrays/s.

serving: a seeded Poisson stream of requests (resolution from {256x256, 640x480, 1280x720,
1920x1080}, spp 1..16, one of 32 resident noise volumes of 64^3..512^3) at 75% of the batching
loop's measured capacity; p50/p99 latency (arrival -> radiance in host memory) for the
continuous-batching loop and for the plain path serving the same arrivals one at a time.

denoiser: a 5-level BF16 conv U-Net trained on the GPU (own kernels) on (1..256 spp, 1024 spp)
pairs of other seeded scenes of the same kind; device time to reach relMSE <= 1e-3 against a
4096-spp render of the benchmark frame, denoised vs plain.

streaming: the terrain world of src/world.rs (16^3 chunks paged around the camera): chunk
upload throughput one add_texture per chunk vs one vt_add_textures batch, a 61-tick camera
flight at the reference's one upload per tick vs draining the queue each tick, and
end_update_instances for an unchanged vs a changed list.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0     # MI355X dense BF16
FRAME_CHUNKS = 3                   # chunks of tile rows (streams) per frame: the library's default
KERNEL_NAMES = {"render": "k_render0", "primary": "k_primary", "secondary": "k_secondary",
                "accumulate": "k_accumulate"}


def algorithmic_bytes(st: dict) -> int:
    return 32 * st["nodes"] + 64 * st["instances"] + 4 * st["bricks"] + 64 * st["bricks_nonempty"] + 5 * st["hits"]


def host_cpus() -> dict:
    """CPU model, nproc, the CPUs this process may run on and its cgroup CPU quota."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    usable = allowed if quota is None else max(1, min(allowed, int(quota)))
    return {"model": model, "nproc": os.cpu_count() or 1, "affinity": allowed, "cgroup_quota": quota,
            "usable": usable}


def native_oracle() -> tuple[str, str]:
    """Builds the oracle with -O3 -march=native on this host (the CPU baseline's build); falls
    back to the -O2 checker build if the compiler is unavailable."""
    odir = os.path.join(ROOT, "oracle")
    try:
        subprocess.run(["make", "-s", "-B", "-C", odir, "native"], check=True, timeout=300, capture_output=True)
        return os.path.join(odir, "_build", "liboracle_native.so"), "-O3 -march=native"
    except (OSError, subprocess.SubprocessError):
        return os.path.join(odir, "_build", "liboracle.so"), "-O2 (native build failed)"


def cpu_baseline(scene, target_s: float = 15.0) -> dict:
    """Times the oracle on whole frames (or every k-th row when a frame would exceed target_s),
    on every CPU this process may use."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O  # the oracle is the CPU baseline's timed code, never the product

    path, flags = native_oracle()
    O.select_library(path)
    cpus = host_cpus()
    threads = cpus["usable"]
    osc = O.OracleScene(scene, use_bvh=True)  # volumes / LBVH prepared outside the timed sample
    H = scene.params.height
    probe = list(range(32, H, 64))
    t0 = time.perf_counter()
    osc.render_rows(probe, threads=threads)
    frame_est = (time.perf_counter() - t0) * H / len(probe)
    if frame_est <= target_s:
        frames = max(1, min(10, int(target_s / max(frame_est, 1e-3))))
        rows = list(range(H))
        what = f"{frames} whole frame(s)"
    else:
        frames = 1
        stride = max(2, int(round(frame_est / target_s)))
        rows = list(range(stride // 2, H, stride))
        what = f"every {stride}th row ({len(rows)} rows) of the frame"
    rays = 0
    t0 = time.perf_counter()
    for _ in range(frames):
        rays += osc.render_rows(rows, threads=threads)["rays"]
    elapsed = time.perf_counter() - t0
    osc.close()
    p = scene.params
    return {"value": round(rays / elapsed / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpus["model"], "nproc": cpus["nproc"], "affinity_cpus": cpus["affinity"],
            "cgroup_cpu_quota": cpus["cgroup_quota"],
            "sample": f"oracle (C++ restatement of the same path, {flags}, {threads} host threads over rows = "
                      f"every CPU this process may use), {what} of the {p.width}x{p.height} {p.spp}-spp "
                      f"workload: {rays} rays in {elapsed:.1f} s"}


def calibrate_copy(nbytes: int = 256 << 20):
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes}), flush=True)


def denoise_frames(args):
    """Only the denoiser: W untimed, then K timed vt_dn_denoise_frame passes (device time each)
    over one frame of the path-tracing scene with its G-buffer; random-init network weights."""
    from vtrace_amd import denoise
    from vtrace_amd import renderer as R
    from vtrace_amd import scenes
    sc = scenes.config2(args.width, args.height, spp=1)
    r = R.get_renderer()
    R.load_scene(r, sc)
    d = denoise.Denoiser(r, seed=5)
    r.enable_gbuffer(True)
    r.render(sc.P, sc.C)
    for _ in range(args.warmup):
        d.denoise_frame()
    ms = [d.denoise_frame()[1] for _ in range(args.denoise_frames)]
    d.close()
    print(json.dumps({"denoise_frame_ms": sorted(ms)[len(ms) // 2], "all_ms": [round(m, 4) for m in ms],
                      "width": sc.params.width, "height": sc.params.height, "params": d.real_param_count}),
          flush=True)


def _kernel_key(name: str):
    """'void (anonymous namespace)::k_secondary<false, false, true>(...)' -> ('k_secondary', False);
    the flag is True for the counting instantiations (first template argument true)."""
    import re
    m = re.search(r"\b(k_\w+|__amd_rocclr_\w+)(<([^>]*)>)?", name)
    if not m:
        return None, False
    args = (m.group(3) or "").split(",")
    return m.group(1), args[0].strip() == "true"


def _counter_rows(d):
    """(kernel name, counter, value, dispatch id) of every counter sample under a rocprofv3
    output directory: the rocpd database (run_results.db) or counter_collection CSV files."""
    import csv
    import glob
    import sqlite3
    for path in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        db = sqlite3.connect(path)
        try:
            for name, counter, value, disp in db.execute(
                    "select kernel_name, counter_name, value, dispatch_id from counters_collection"):
                yield name, counter, float(value), disp
        finally:
            db.close()
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(path)):
            yield (row.get("Kernel_Name", ""), row["Counter_Name"], float(row["Counter_Value"]),
                   row.get("Dispatch_Id") or row.get("Correlation_Id"))


def summarize_counters(dirs) -> dict:
    """Mean per dispatch of every counter of every kernel in rocprofv3 --pmc outputs,
    counting-frame instantiations left out."""
    sums, disp = {}, {}
    for d in dirs:
        for name, c, value, dispatch in _counter_rows(d):
            k, counting = _kernel_key(name)
            if k is None or counting:
                continue
            sums.setdefault(k, {}).setdefault(c, 0.0)
            sums[k][c] += value
            disp.setdefault(k, {}).setdefault(c, set()).add(dispatch)
    return {k: {c: v / max(1, len(disp[k][c])) for c, v in cs.items()} for k, cs in sums.items()}


def summarize_trace(d) -> list:
    """[(kernel, calls, average us, min us, max us, total ms)] from a rocprofv3 --kernel-trace
    database, slowest total first."""
    import glob
    import sqlite3
    out = {}
    for path in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        db = sqlite3.connect(path)
        try:
            for name, dur in db.execute("select name, duration from kernels"):
                k = name.replace("(anonymous namespace)::", "").replace("void ", "", 1).split("(")[0]
                out.setdefault(k, []).append(float(dur) * 1e-3)
        finally:
            db.close()
    rows = [(k, len(v), sum(v) / len(v), min(v), max(v), sum(v) * 1e-3) for k, v in out.items()]
    return sorted(rows, key=lambda r: -r[5])


def _dispatch_rows(d):
    """Per-dispatch {name, counters...} in dispatch order from a rocprofv3 --pmc directory."""
    rows = {}
    for name, c, v, disp in _counter_rows(d):
        e = rows.setdefault(disp, {"name": name})
        e[c] = e.get(c, 0.0) + v
    return [rows[k] for k in sorted(rows, key=lambda x: int(x))]


def _short(name: str) -> str:
    return name.replace("(anonymous namespace)::", "").replace("void ", "", 1).split("(")[0]


def summarize_launches(args) -> None:
    """Markdown table of every launch of the last frame (from the last launch of --frame-from)
    of a rocprofv3 --kernel-trace directory: duration, and where the PMC directories are given,
    BF16 MFMA TFLOP/s and its fraction of peak, MFMA busy (share of SIMD cycles) and HBM bytes
    (FETCH_SIZE / WRITE_SIZE scaled by a plain copy of known size)."""
    import glob
    import sqlite3
    trace = []
    for path in glob.glob(os.path.join(args.trace, "**", "*.db"), recursive=True):
        db = sqlite3.connect(path)
        try:
            trace += list(db.execute("select name, duration, start from kernels"))
        finally:
            db.close()
    trace.sort(key=lambda r: r[2])

    def last_frame(rows, key):
        idx = [i for i, r in enumerate(rows) if args.frame_from in key(r)]
        return rows[idx[-1]:] if idx else []

    launches = last_frame(trace, lambda r: r[0])
    mfma = last_frame(_dispatch_rows(args.mfma), lambda r: r["name"]) if args.mfma else []
    fetch = last_frame(_dispatch_rows(args.fetch), lambda r: r["name"]) if args.fetch else []
    write = last_frame(_dispatch_rows(args.write), lambda r: r["name"]) if args.write else []
    fk = wk = None
    if args.cal_fetch and args.cal_write:
        copy = "__amd_rocclr_copyBuffer"
        fk = summarize_counters([args.cal_fetch])[copy]["FETCH_SIZE"] * 1024.0 / (256 << 20)
        wk = summarize_counters([args.cal_write])[copy]["WRITE_SIZE"] * 1024.0 / (256 << 20)
    print("| kernel | us | BF16 TFLOP/s | of peak | MFMA busy | fetch MB | write MB |\n|---|---|---|---|---|---|---|")
    tot_us = tot_fl = tot_f = tot_w = 0.0
    for i, (name, dur, _) in enumerate(launches):
        us = float(dur) * 1e-3
        tot_us += us
        cells = [f"`{_short(name)}`", f"{us:.1f}"]
        m = mfma[i] if i < len(mfma) and _short(mfma[i]["name"]) == _short(name) else None
        if m and m.get("SQ_INSTS_VALU_MFMA_MOPS_BF16"):
            fl = m["SQ_INSTS_VALU_MFMA_MOPS_BF16"] * 512.0
            tot_fl += fl
            tf = fl / (us * 1e-6) / 1e12
            # SQ_VALU_MFMA_BUSY_CYCLES summed over SIMDs; GRBM_GUI_ACTIVE summed over 8 XCDs
            busy = m["SQ_VALU_MFMA_BUSY_CYCLES"] / (m["GRBM_GUI_ACTIVE"] / 8.0 * 1024.0) if m.get("GRBM_GUI_ACTIVE") else None
            cells += [f"{tf:.0f}", f"{tf / BF16_MFMA_PEAK_TFLOPS:.3f}", f"{busy:.3f}" if busy is not None else ""]
        else:
            cells += ["", "", ""]
        f = fetch[i] if i < len(fetch) and _short(fetch[i]["name"]) == _short(name) else None
        w = write[i] if i < len(write) and _short(write[i]["name"]) == _short(name) else None
        if f and w and fk and wk:
            fb, wb = f["FETCH_SIZE"] * 1024.0 / fk, w["WRITE_SIZE"] * 1024.0 / wk
            tot_f += fb
            tot_w += wb
            cells += [f"{fb / 1e6:.1f}", f"{wb / 1e6:.1f}"]
        else:
            cells += ["", ""]
        print("| " + " | ".join(cells) + " |")
    tf = tot_fl / (tot_us * 1e-6) / 1e12 if tot_us else 0.0
    print(f"| **frame** | {tot_us:.1f} | {tf:.0f} | {tf / BF16_MFMA_PEAK_TFLOPS:.3f} | | {tot_f / 1e6:.1f} | {tot_w / 1e6:.1f} |")


def summarize_profiles(args) -> None:
    """Writes profiles/traffic.json (FETCH_SIZE / WRITE_SIZE per launch, scaled by a plain copy
    of known size) and profiles/pmc.json (SQ / TCC / TCP counters per launch) from rocprofv3
    output directories."""
    cmd = " ".join(sys.argv)
    if args.fetch and args.write:
        fetch, write = summarize_counters([args.fetch]), summarize_counters([args.write])
        cf, cw = summarize_counters([args.cal_fetch]), summarize_counters([args.cal_write])
        copy = "__amd_rocclr_copyBuffer"
        fk = cf[copy]["FETCH_SIZE"] * 1024.0 / (256 << 20)
        wk = cw[copy]["WRITE_SIZE"] * 1024.0 / (256 << 20)
        kernels = {}
        for k in sorted(set(fetch) & set(write)):
            kernels[k] = {"fetch_bytes_per_launch": int(fetch[k]["FETCH_SIZE"] * 1024.0 / fk),
                          "write_bytes_per_launch": int(write[k]["WRITE_SIZE"] * 1024.0 / wk)}
        out = {"workload": args.workload, "command": cmd,
               "calibration": {"copy_bytes": 256 << 20, "fetch_factor": round(fk, 4), "write_factor": round(wk, 4),
                               "copy_kernel": copy},
               "kernels": kernels}
        json.dump(out, open(os.path.join(ROOT, "profiles", "traffic.json"), "w"), indent=1)
    if args.pmc:
        out = {"workload": args.workload, "command": cmd, "kernels": summarize_counters(args.pmc)}
        json.dump(out, open(os.path.join(ROOT, "profiles", "pmc.json"), "w"), indent=1)
    if args.trace:
        print("| kernel | calls | avg us | min us | max us | total ms |\n|---|---|---|---|---|---|")
        for k, n, avg, mn, mx, tot in summarize_trace(args.trace):
            print(f"| `{k}` | {n} | {avg:.1f} | {mn:.1f} | {mx:.1f} | {tot:.3f} |")


def load_traffic(kernel: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        t = json.load(open(path))
        k = t["kernels"][kernel]
        return int(k["fetch_bytes_per_launch"] + k["write_bytes_per_launch"])
    except (KeyError, ValueError, TypeError):
        return None


def load_pmc(kernel: str, workload: str):
    """Lane utilisation, VALU busy and L2 hit rate of a kernel from profiles/pmc.json (None if
    the profile is missing or was taken on another workload)."""
    path = os.path.join(ROOT, "profiles", "pmc.json")
    if not os.path.exists(path):
        return None
    try:
        t = json.load(open(path))
        if t.get("workload") != workload:
            return None
        c = t["kernels"][kernel]
        out = {}
        if c.get("SQ_ACTIVE_INST_VALU"):
            out["lane_utilisation"] = round(c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"]), 4)
        if "VALUBusy" in c:
            out["valu_busy_pct"] = round(c["VALUBusy"], 1)
        if c.get("TCC_HIT_sum") is not None and c.get("TCC_MISS_sum") is not None:
            out["l2_hit"] = round(c["TCC_HIT_sum"] / max(1.0, c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4)
        return out or None
    except (KeyError, ValueError, TypeError, ZeroDivisionError):
        return None


def limiter(hbm_frac_measured, pmc) -> str:
    """What binds the kernel, from its measured HBM bandwidth and lane utilisation."""
    if hbm_frac_measured is not None and hbm_frac_measured >= 0.5:
        return "hbm"
    if pmc and pmc.get("lane_utilisation") is not None and pmc["lane_utilisation"] < 0.5 and \
            (hbm_frac_measured is None or hbm_frac_measured < 0.2):
        return "valu-divergence"
    if pmc and pmc.get("valu_busy_pct", 0) >= 70:
        return "valu"
    return "latency"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only a plain copy kernel (PMC scaling)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests in the serving stream (0 skips the serving measurement)")
    ap.add_argument("--denoiser-iters", type=int, default=24000,
                    help="training iterations of the denoiser (0 skips the denoiser measurement)")
    ap.add_argument("--streaming", type=int, default=1, help="1: streamed terrain measurement, 0: skip it")
    ap.add_argument("--denoise-frames", type=int, default=0,
                    help="run only the denoiser: this many 1920x1080 inference passes over a frame of the "
                         "path-tracing scene (random-init network), for profiling")
    ap.add_argument("--single-stream", action="store_true",
                    help="render every frame on one stream (kernel traces then match the stage table)")
    ap.add_argument("--summarize-profiles", action="store_true",
                    help="write profiles/traffic.json and profiles/pmc.json from rocprofv3 output directories")
    ap.add_argument("--workload", default="c2_16x256cube_1920x1080_4spp_4b")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cal-fetch")
    ap.add_argument("--cal-write")
    ap.add_argument("--pmc", nargs="*")
    ap.add_argument("--trace", help="print a markdown table of a rocprofv3 --kernel-trace output directory")
    ap.add_argument("--frame-from", help="with --trace: one row per launch of the last frame, which starts at the "
                                         "last launch of this kernel (e.g. k_features)")
    ap.add_argument("--mfma", help="with --frame-from: rocprofv3 --pmc directory with SQ_INSTS_VALU_MFMA_MOPS_BF16, "
                                   "SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE")
    args = ap.parse_args()
    if args.summarize_profiles:
        if args.frame_from:
            summarize_launches(args)
        else:
            summarize_profiles(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the renderer has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.denoise_frames > 0:
        denoise_frames(args)
        return

    from vtrace_amd import renderer as R
    from vtrace_amd import scenes

    t0 = time.perf_counter()
    sc = scenes.config2(args.width, args.height)
    t_gen = time.perf_counter() - t0
    r = R.get_renderer()
    r.lib.vt_set_frame_split(1 if args.single_stream else FRAME_CHUNKS)
    t0 = time.perf_counter()
    R.load_scene(r, sc, stats=True)
    t_up = time.perf_counter() - t0
    print(f"# scene {sc.name}: generated in {t_gen:.1f} s, uploaded+bricked in {t_up:.1f} s", file=sys.stderr,
          flush=True)
    # counting frame (untimed): rays and algorithmic bytes per launch stage
    r.render(sc.P, sc.C)
    st = r.stats()
    stages = r.stage_stats()
    pixels = sc.params.width * sc.params.height
    hits = r.hits()
    hit_frac = float((hits["inst"] != 0xFFFFFFFF).mean())
    r.set_params(sc.params, stats=False)
    r.render(sc.P, sc.C)
    # timed: W untimed frames, then exactly K frames as render_tick renders them (each complete
    # before the next); wall_ms = host time per frame, kernel_ms = GPU time per frame
    r.lib.vt_set_frame_split(1 if args.single_stream else FRAME_CHUNKS)
    r.bench(args.warmup, 1) if args.warmup > 0 else None
    r.lib.vt_synchronize()
    t0 = time.perf_counter()
    kernel_ms, wall_ms = r.bench(0, args.steps)
    r.lib.vt_synchronize()
    t_all = time.perf_counter() - t0
    # launch-stage breakdown (untimed for the metric): the same frames on one stream, HIP events
    # around every launch; the roofline of the dominant kernel comes from these
    r.lib.vt_set_frame_split(1)
    single_kernel_ms, single_wall_ms, stage_ms = r.bench_stages(args.warmup, args.steps)
    r.lib.vt_set_frame_split(1 if args.single_stream else FRAME_CHUNKS)
    # per-frame latency distribution (render_tick round trips, GPU time)
    lat = []
    for _ in range(max(5, min(args.steps, 20))):
        r.render(sc.P, sc.C)
        lat.append(r.last_frame_ms())
    lat.sort()
    p50 = lat[len(lat) // 2]
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]

    per_kind = {}
    for (kind, s), (kind2, ms) in zip(stages, stage_ms):
        assert kind == kind2
        b = algorithmic_bytes(s)
        if kind == "primary":
            b += 16 * pixels
        if kind == "accumulate":
            b += 16 * pixels
        e = per_kind.setdefault(kind, {"launches": 0, "bytes": 0, "ms": 0.0, "rays": 0, "counters": {}, "launch_ms": []})
        e["launches"] += 1
        e["bytes"] += b
        e["ms"] += ms
        e["rays"] += s["rays"]
        e["launch_ms"].append(round(ms, 4))
        for c, v in s.items():
            e["counters"][c] = e["counters"].get(c, 0) + v
    dom = max(per_kind, key=lambda k: per_kind[k]["ms"])
    d = per_kind[dom]
    launch_ms = d["ms"] / d["launches"]
    bytes_launch = d["bytes"] / d["launches"]
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9
    frame_bytes = sum(e["bytes"] for e in per_kind.values())
    mrays = st["rays"] / (wall_ms * 1e-3) / 1e6
    kname = KERNEL_NAMES[dom]
    traffic = load_traffic(kname)
    hbm_measured = traffic / (launch_ms * 1e-3) / 1e9 if traffic else None
    pmc = load_pmc(kname, sc.name)
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
        "roofline": {"bound": limiter(hbm_measured / HBM_PEAK_GBS if hbm_measured else None, pmc),
                     "roof": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "hbm_gbs_measured": round(hbm_measured, 1) if hbm_measured else None,
                     "pmc": pmc,
                     "kernel": kname, "launches_per_frame": d["launches"], "launch_ms": round(launch_ms, 4),
                     "algorithmic_bytes_per_launch": int(bytes_launch),
                     "frame_algorithmic_bytes": int(frame_bytes),
                     "frame_achieved_gbs": round(frame_bytes / (kernel_ms * 1e-3) / 1e9, 1),
                     "bytes_per_ray": round(frame_bytes / max(1, st["rays"]), 1)},
        "stages": {k: {"launches": v["launches"], "ms": round(v["ms"], 4), "rays": v["rays"],
                       "GBs": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1), "launch_ms": v["launch_ms"],
                       "counters": v["counters"]} for k, v in per_kind.items()},
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4)},
        "kernel_ms_per_frame": round(kernel_ms, 4),
        "single_stream_ms_per_frame": round(single_wall_ms, 4),
        "counters": st,
        "timed_region_s": round(t_all, 3),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc)
    if args.serving_requests > 0:
        from vtrace_amd import serving
        out["serving"] = serving.benchmark(r, nreq=args.serving_requests,
                                           log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    if args.denoiser_iters > 0:
        from vtrace_amd import denoise
        out["denoiser"] = denoise.benchmark(r, sc, iters=args.denoiser_iters,
                                            log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    if args.streaming:
        from vtrace_amd import world
        out["streaming"] = world.benchmark(r, args.width, args.height,
                                           log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
