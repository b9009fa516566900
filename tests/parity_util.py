"""Helpers shared by the GPU parity tests: run a scene on the HIP path and on the oracle."""
from __future__ import annotations

import copy

import numpy as np

import oracle_ffi as O
from vtrace_amd import renderer as R

HIT_FIELDS = ("inst", "voxel", "steps", "face", "pal", "mask", "flags", "vhi")


def gpu_render(scene, stats=False, params=None):
    r = R.reset_renderer()
    sc = scene if params is None else copy.copy(scene)
    if params is not None:
        sc.params = params
    R.load_scene(r, sc, stats=stats)
    r.render(sc.P, sc.C)
    hits = r.hits()
    rad = r.radiance()
    st = r.stats() if stats else None
    return hits, rad, st


def assert_hits_equal(gpu, orc, label=""):
    assert gpu.shape == orc.shape
    for f in HIT_FIELDS:
        bad = np.argwhere(gpu[f] != orc[f])
        if len(bad):
            y, x = bad[0]
            raise AssertionError(
                f"{label}: field {f} differs at {len(bad)} pixels; first (y={y}, x={x}): "
                f"gpu={gpu[y, x]} oracle={orc[y, x]}")


def radiance_error(gpu, orc):
    """Per-pixel relative error of the RGB radiance (FP32 GPU vs FP64 oracle)."""
    g = gpu[..., :3].astype(np.float64)
    o = orc[..., :3].astype(np.float64)
    err = np.abs(g - o) / np.maximum(np.abs(o), 1e-6)
    return float(err.max()), float(err.mean())
