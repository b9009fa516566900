"""Loader for the native libraries of vtrace_amd.

``librender_hip.so`` is the HIP drop-in for the reference's C library (the exports
``src/render.rs:110-128`` binds over FFI, declared in ``include/vtrace.h``).
``libvtscene.so`` is the host-side synthetic scene generator.

Both are built in-tree by ``__graft_entry__.build()``.  There is no fallback: if a
library is missing, loading raises, so nothing can silently run on a CPU path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
RENDER_LIB = os.path.join(LIB_DIR, "librender_hip.so")
SCENE_LIB = os.path.join(LIB_DIR, "libvtscene.so")

_cache: dict[str, ctypes.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str) -> ctypes.CDLL:
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[path] = lib
    return lib


def scene_lib() -> ctypes.CDLL:
    lib = _load(SCENE_LIB)
    if not getattr(lib, "_vt_typed", False):
        c = ctypes
        lib.vt_pcg_hash.restype = c.c_uint32
        lib.vt_pcg_hash.argtypes = [c.c_uint32]
        lib.vt_gen_noise_field.restype = c.c_int
        lib.vt_gen_noise_field.argtypes = [c.c_uint32] * 5 + [c.c_void_p]
        lib.vt_gen_volume.restype = c.c_int
        lib.vt_gen_volume.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32, c.c_double,
                                      c.c_uint32, c.c_uint32, c.c_void_p, c.c_void_p]
        lib.vt_gen_palette.restype = None
        lib.vt_gen_palette.argtypes = [c.c_uint32, c.c_int, c.c_void_p]
        lib.vt_expand_palette.restype = None
        lib.vt_expand_palette.argtypes = [c.c_void_p, c.c_size_t, c.c_void_p, c.c_void_p]
        lib.vt_vox_models.restype = c.c_int32
        lib.vt_vox_models.argtypes = [c.c_void_p, c.c_uint64]
        lib.vt_vox_model.restype = c.c_int32
        lib.vt_vox_model.argtypes = [c.c_void_p, c.c_uint64, c.c_uint32, c.c_void_p, c.c_void_p, c.c_void_p]
        lib.vt_terrain_voxel.restype = c.c_uint32
        lib.vt_terrain_voxel.argtypes = [c.c_int32, c.c_int32, c.c_int32, c.c_uint32]
        lib.vt_terrain_chunk.restype = c.c_int32
        lib.vt_terrain_chunk.argtypes = [c.c_int32, c.c_int32, c.c_int32, c.c_uint32, c.c_void_p]
        lib.vt_terrain_chunks.restype = c.c_int32
        lib.vt_terrain_chunks.argtypes = [c.c_void_p, c.c_uint32, c.c_uint32, c.c_void_p, c.c_void_p]
        lib._vt_typed = True
    return lib


def render_lib() -> ctypes.CDLL:
    from . import abi  # noqa: F401  (declares argtypes on first load)
    lib = _load(RENDER_LIB)
    abi.declare(lib)
    return lib
