"""MagicaVoxel .vox assets as volumes for the renderer (the reference's asset path,
src/voxel/magica_voxel.rs:18-44, render.rs:246-262).

Parsing is native (vt_vox_* in libvtscene, vox.cpp).  Each model becomes a Volume whose u8
index array is the byte layout the reference hands to add_texture, with a palette that
expands it to the reference's RGBA8 texels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import scenes
from ._native import scene_lib


class VoxError(ValueError):
    pass


_ERRORS = {-1: "malformed .vox data or no such model",
           -2: "a voxel falls outside its chunk (the model is not a cube)",
           -3: "no RGBA palette chunk (the format's default palette is not supported)"}


def parse(data: bytes) -> list:
    """All models of a .vox buffer as scenes.Volume (u8 indices + palette)."""
    lib = scene_lib()
    buf = np.frombuffer(data, dtype=np.uint8)
    n = lib.vt_vox_models(buf.ctypes.data, buf.size)
    if n < 0:
        raise VoxError(_ERRORS[-1])
    out = []
    for k in range(n):
        dims = np.zeros(3, dtype=np.uint32)
        rc = lib.vt_vox_model(buf.ctypes.data, buf.size, k, dims.ctypes.data, None, None)
        if rc != 0:
            raise VoxError(_ERRORS.get(rc, "unknown error"))
        dx, dy, dz = (int(v) for v in dims)
        idx = np.zeros((dz, dy, dx), dtype=np.uint8)
        pal = np.zeros(256, dtype=np.uint32)
        rc = lib.vt_vox_model(buf.ctypes.data, buf.size, k, dims.ctypes.data, idx.ctypes.data, pal.ctypes.data)
        if rc != 0:
            raise VoxError(_ERRORS.get(rc, "unknown error"))
        out.append(scenes.Volume(idx, pal))
    return out


def load(path: str) -> list:
    with open(path, "rb") as f:
        return parse(f.read())
