"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module.  The oracle is the checker, never the thing measured.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class OrcVolume(C.Structure):
    _fields_ = [("dim", C.c_uint32 * 3), ("fmt", C.c_uint32), ("data", C.c_void_p), ("palette", C.c_void_p)]


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("viewport_square", C.c_uint32),
                ("mode", C.c_uint32), ("spp", C.c_uint32), ("max_bounces", C.c_uint32),
                ("jitter", C.c_uint32), ("seed", C.c_uint32), ("frame", C.c_uint32),
                ("sun_dir", C.c_float * 3), ("sun_radiance", C.c_float * 3), ("sky", C.c_float * 3)]


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "nodes", "instances", "bricks", "bricks_nonempty",
                                          "hits", "steps")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


HIT_DTYPE = np.dtype([("inst", "<u4"), ("voxel", "<u4"), ("steps", "<u2"), ("face", "u1"), ("pal", "u1"),
                      ("mask", "u1"), ("flags", "u1"), ("vhi", "<u2")])
assert HIT_DTYPE.itemsize == 16

_lib = None


def select_library(path: str) -> None:
    """Load the oracle from `path` (another build of oracle.cpp, e.g. the -O3 -march=native
    build the CPU baseline times) instead of the checker build.  Must precede the first lib()."""
    global ORACLE_LIB
    if _lib is not None and ORACLE_LIB != path:
        raise RuntimeError("the oracle library is already loaded from " + ORACLE_LIB)
    ORACLE_LIB = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_LIB)
        L.orc_pcg_hash.restype = C.c_uint32
        L.orc_pcg_hash.argtypes = [C.c_uint32]
        L.orc_rng_seed.restype = C.c_uint32
        L.orc_rng_seed.argtypes = [C.c_uint32] * 4
        L.orc_morton30.restype = C.c_uint32
        L.orc_morton30.argtypes = [C.c_uint32] * 3
        L.orc_inverse4.restype = C.c_int
        L.orc_inverse4.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_srgb_decode.restype = C.c_float
        L.orc_srgb_decode.argtypes = [C.c_uint32]
        L.orc_srgb_encode.restype = C.c_uint32
        L.orc_srgb_encode.argtypes = [C.c_float]
        L.orc_nearest_texel.restype = C.c_uint32
        L.orc_nearest_texel.argtypes = [C.c_int, C.c_uint32]
        L.orc_brick_blocks.restype = None
        L.orc_brick_blocks.argtypes = [C.POINTER(OrcVolume), C.c_void_p]
        L.orc_cosine_dir_samples.restype = None
        L.orc_cosine_dir_samples.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_instance_prep.restype = None
        L.orc_instance_prep.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        L.orc_lbvh_build.restype = C.c_int
        L.orc_lbvh_build.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_brick.restype = C.c_uint32
        L.orc_brick.argtypes = [C.POINTER(OrcVolume), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_render.restype = C.c_int
        L.orc_render.argtypes = [C.c_uint32, C.POINTER(OrcVolume), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(OrcParams), C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.POINTER(OrcStats)]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_uint32, C.POINTER(OrcVolume), C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(OrcParams), C.c_int]
        L.orc_scene_destroy.restype = None
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_render.restype = C.c_int
        L.orc_scene_render.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(OrcStats)]
        L.orc_scene_render_rows.restype = C.c_int
        L.orc_scene_render_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.POINTER(OrcStats)]
        L.orc_dda_test.restype = C.c_int
        L.orc_dda_test.argtypes = [C.POINTER(OrcVolume), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        _lib = L
    return _lib


def _vol_structs(volumes):
    arr = (OrcVolume * max(1, len(volumes)))()
    keep = []
    for i, v in enumerate(volumes):
        data = np.ascontiguousarray(v.data)
        keep.append(data)
        arr[i].dim[0], arr[i].dim[1], arr[i].dim[2] = v.dims
        arr[i].fmt = v.fmt
        arr[i].data = data.ctypes.data
        if v.fmt == 0:
            pal = np.ascontiguousarray(v.palette, dtype=np.uint32)
            keep.append(pal)
            arr[i].palette = pal.ctypes.data
        else:
            arr[i].palette = None
    return arr, keep


def params_struct(p) -> OrcParams:
    s = OrcParams()
    s.width, s.height, s.viewport_square = p.width, p.height, p.viewport_square
    s.mode, s.spp, s.max_bounces, s.jitter = p.mode, p.spp, p.max_bounces, p.jitter
    s.seed, s.frame = p.seed & 0xFFFFFFFF, p.frame & 0xFFFFFFFF
    for k in range(3):
        s.sun_dir[k] = p.sun_dir[k]
        s.sun_radiance[k] = p.sun_radiance[k]
        s.sky[k] = p.sky[k]
    return s


def render(scene, use_bvh: bool = False, rows=None, params=None):
    """Run the oracle over a Scene.  Returns (hits, radiance, stats).

    radiance is float32 (W,H,4) for mode 0 and float64 for modes 1/2."""
    p = params or scene.params
    W, H = p.width, p.height
    y0, y1 = (0, H) if rows is None else rows
    vols, keep = _vol_structs(scene.volumes)
    recs = np.ascontiguousarray(scene.records, dtype=np.float32)
    hits = np.zeros(W * H, dtype=HIT_DTYPE)
    rad32 = np.zeros((H, W, 4), dtype=np.float32)
    rad64 = np.zeros((H, W, 4), dtype=np.float64)
    st = OrcStats()
    ps = params_struct(p)
    P = np.ascontiguousarray(scene.P, dtype=np.float32)
    Cm = np.ascontiguousarray(scene.C, dtype=np.float32)
    rc = lib().orc_render(len(scene.volumes), vols, recs.shape[0], recs.ctypes.data, P.ctypes.data,
                          Cm.ctypes.data, C.byref(ps), 1 if use_bvh else 0, y0, y1, hits.ctypes.data,
                          rad32.ctypes.data, rad64.ctypes.data, C.byref(st))
    if rc != 0:
        raise RuntimeError("orc_render failed")
    del keep
    return hits.reshape(H, W), (rad32 if p.mode == 0 else rad64), st.as_dict()


def brick(volume):
    vols, keep = _vol_structs([volume])
    dx, dy, dz = volume.dims
    nb = ((dx + 7) // 8) * ((dy + 7) // 8) * ((dz + 7) // 8)
    bmap = np.empty(nb, dtype=np.uint32)
    count = lib().orc_brick(vols, bmap.ctypes.data, None, None)
    masks = np.empty((max(count, 1), 8), dtype=np.uint64)
    bpv = 1 if volume.fmt == 0 else 4
    payload = np.empty(max(count, 1) * 512 * bpv, dtype=np.uint8)
    lib().orc_brick(vols, bmap.ctypes.data, masks.ctypes.data, payload.ctypes.data)
    return bmap, masks[:count], payload[: count * 512 * bpv]


def lbvh(records, nvol):
    recs = np.ascontiguousarray(records, dtype=np.float32)
    n = recs.shape[0]
    nodes = np.zeros((max(1, 2 * n - 1), 8), dtype=np.uint32)
    order = np.zeros(max(1, n), dtype=np.uint32)
    codes = np.zeros(max(1, n), dtype=np.uint32)
    lib().orc_lbvh_build(n, recs.ctypes.data, nvol, nodes.ctypes.data, order.ctypes.data, codes.ctypes.data)
    return nodes[: max(0, 2 * n - 1)], order[:n], codes[:n]


def instance_prep(records, nvol):
    recs = np.ascontiguousarray(records, dtype=np.float32)
    n = recs.shape[0]
    inv = np.zeros((n, 16), dtype=np.float32)
    aabb = np.zeros((n, 6), dtype=np.float32)
    tex = np.zeros(n, dtype=np.uint32)
    valid = np.zeros(n, dtype=np.uint8)
    lib().orc_instance_prep(n, recs.ctypes.data, nvol, inv.ctypes.data, aabb.ctypes.data, tex.ctypes.data,
                            valid.ctypes.data)
    return inv, aabb, tex, valid


class OracleScene:
    """A scene prepared once in the oracle (volumes, instances, LBVH); rows rendered on demand."""

    def __init__(self, scene, use_bvh: bool = True, params=None):
        self.p = params or scene.params
        self._vols, self._keep = _vol_structs(scene.volumes)
        self._recs = np.ascontiguousarray(scene.records, dtype=np.float32)
        self._P = np.ascontiguousarray(scene.P, dtype=np.float32)
        self._C = np.ascontiguousarray(scene.C, dtype=np.float32)
        self._ps = params_struct(self.p)
        self.h = lib().orc_scene_create(len(scene.volumes), self._vols, self._recs.shape[0], self._recs.ctypes.data,
                                        self._P.ctypes.data, self._C.ctypes.data, C.byref(self._ps),
                                        1 if use_bvh else 0)
        if not self.h:
            raise RuntimeError("orc_scene_create failed")

    def render(self, rows=None, threads: int = 1):
        W, H = self.p.width, self.p.height
        y0, y1 = (0, H) if rows is None else rows
        hits = np.zeros(W * H, dtype=HIT_DTYPE)
        rad32 = np.zeros((H, W, 4), dtype=np.float32)
        rad64 = np.zeros((H, W, 4), dtype=np.float64)
        st = OrcStats()
        rc = lib().orc_scene_render(self.h, y0, y1, threads, hits.ctypes.data, rad32.ctypes.data, rad64.ctypes.data,
                                    C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_scene_render failed")
        return hits.reshape(H, W), (rad32 if self.p.mode == 0 else rad64), st.as_dict()

    def render_rows(self, rows, threads: int = 1):
        """Renders the listed rows (counters only; outputs of those rows are discarded)."""
        W, H = self.p.width, self.p.height
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        hits = np.zeros(W * H, dtype=HIT_DTYPE)
        rad32 = np.zeros((H, W, 4), dtype=np.float32)
        rad64 = np.zeros((H, W, 4), dtype=np.float64)
        st = OrcStats()
        rc = lib().orc_scene_render_rows(self.h, r.ctypes.data, len(r), threads, hits.ctypes.data, rad32.ctypes.data,
                                         rad64.ctypes.data, C.byref(st))
        if rc != 0:
            raise RuntimeError("orc_scene_render_rows failed")
        return st.as_dict()

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def dda(volume, level, mpos, v0, mdir):
    """Oracle DDA on one volume: level 0 = reference DDA, 1 = two-level secondary DDA.
    Returns (hit, (x, y, z), steps, mask, t_last, q)."""
    vols, keep = _vol_structs([volume])
    mp = np.ascontiguousarray(mpos, dtype=np.float32)
    vv = np.ascontiguousarray(v0, dtype=np.int32)
    md = np.ascontiguousarray(mdir, dtype=np.float32)
    out = np.zeros(6, dtype=np.int32)
    tq = np.zeros(4, dtype=np.float32)
    lib().orc_dda_test(vols, level, mp.ctypes.data, vv.ctypes.data, md.ctypes.data, out.ctypes.data, tq.ctypes.data)
    return bool(out[0]), (int(out[1]), int(out[2]), int(out[3])), int(out[4]), int(out[5]), float(tq[0]), tq[1:].copy()


def cosine_dir_samples(seed: int, normal, n: int) -> np.ndarray:
    """n bounce directions about a unit normal from the oracle's cosine sampler (n x 3 float32)."""
    nrm = np.ascontiguousarray(normal, dtype=np.float32)
    out = np.empty((n, 3), dtype=np.float32)
    lib().orc_cosine_dir_samples(seed, nrm.ctypes.data, n, out.ctypes.data)
    return out


def brick_blocks(volume) -> np.ndarray:
    """2^3-block masks of every brick of a volume (brick order, 0 for empty bricks)."""
    vols, keep = _vol_structs([volume])
    bd = [(d + 7) // 8 for d in volume.dims]
    out = np.zeros(bd[0] * bd[1] * bd[2], dtype=np.uint64)
    lib().orc_brick_blocks(vols, out.ctypes.data)
    _ = keep
    return out
