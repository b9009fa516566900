"""Seeded scenes for the renderer: volumes, palettes, instances and cameras.

Camera and instance conventions follow the reference host code:
* perspective: GL-style, fov 80 degrees, near 0.01, far 10000 (``src/render.rs:190,208-210``);
* camera: ``look_at(eye, centre, up=(0,1,0))`` (``src/render.rs:212-214``);
* instance record: one column-major mat4 with the texture id bit-cast into ``m[3][3]``
  (``src/render.rs:31-35,74-78``, ``shaders/trace.vert:37-40``);
* one world unit per voxel: ``model = translate(t) * scale(N)``.

Matrices are computed in float64 and handed over as float32 bytes, exactly what
crosses the reference's FFI (``render_tick_info``, ``lib/common.h:89-92``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._native import scene_lib

FOV = 80.0 / 180.0 * 3.1415926          # src/render.rs:190
NEAR, FAR = 0.01, 10000.0               # src/render.rs:209
SKY = (0.53, 0.81, 0.92)                # lib/command.c:57-60 clear colour


def pcg_hash(v: int) -> int:
    return int(scene_lib().vt_pcg_hash(v & 0xFFFFFFFF))


class Rng:
    """Small seeded stream (PCG hash chain) for scene parameters."""

    def __init__(self, seed: int):
        self.s = pcg_hash(seed ^ 0x2545F491)

    def u32(self) -> int:
        self.s = pcg_hash(self.s)
        return self.s

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.u32() >> 8) / 16777216.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.u32() % (hi - lo + 1)


def perspective(fovy: float, aspect: float, near: float = NEAR, far: float = FAR) -> np.ndarray:
    """GL right-handed perspective, column-major float32[16] (m[c*4 + r])."""
    f = 1.0 / math.tan(fovy / 2.0)
    m = np.zeros((4, 4))  # m[c][r]
    m[0][0] = f / aspect
    m[1][1] = f
    m[2][2] = -(far + near) / (far - near)
    m[2][3] = -1.0
    m[3][2] = -(2.0 * far * near) / (far - near)
    return m.astype(np.float32).reshape(16)


def look_at(eye, centre, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """GL right-handed look-at, column-major float32[16]."""
    eye, centre, up = (np.asarray(v, dtype=np.float64) for v in (eye, centre, up))
    f = centre - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.zeros((4, 4))
    m[0][0], m[1][0], m[2][0] = s
    m[0][1], m[1][1], m[2][1] = u
    m[0][2], m[1][2], m[2][2] = -f
    m[3][0] = -np.dot(s, eye)
    m[3][1] = -np.dot(u, eye)
    m[3][2] = np.dot(f, eye)
    m[3][3] = 1.0
    return m.astype(np.float32).reshape(16)


def model_matrix(translate, scale) -> np.ndarray:
    """translate(t) * scale(s) as column-major float32[16] with m[3][3] = 1."""
    m = np.zeros((4, 4))
    sx, sy, sz = (scale, scale, scale) if np.isscalar(scale) else scale
    m[0][0], m[1][1], m[2][2] = sx, sy, sz
    m[3][0], m[3][1], m[3][2] = translate
    m[3][3] = 1.0
    return m.astype(np.float32).reshape(16)


def instance_record(model: np.ndarray, texture_id: int) -> np.ndarray:
    """GPUInstance::from_model (src/render.rs:74-78): texture id bits in m[3][3]."""
    rec = np.array(model, dtype=np.float32).copy()
    rec.view(np.uint32)[15] = np.uint32(texture_id)
    return rec


def gen_volume(dims, seed: int, fill: float, octaves: int = 4, boxes=None) -> np.ndarray:
    """u8 palette-index volume, x fastest (shape (dz, dy, dx))."""
    dx, dy, dz = dims
    out = np.empty(dx * dy * dz, dtype=np.uint8)
    if boxes is not None and len(boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.uint32).reshape(-1, 6))
        nb, bp = b.shape[0], b.ctypes.data
    else:
        nb, bp = 0, None
    rc = scene_lib().vt_gen_volume(dx, dy, dz, seed, float(fill), octaves, nb, bp, out.ctypes.data)
    if rc != 0:
        raise ValueError("vt_gen_volume failed")
    return out.reshape(dz, dy, dx)


def gen_palette(seed: int, opaque: bool = True) -> np.ndarray:
    out = np.empty(256, dtype=np.uint32)
    scene_lib().vt_gen_palette(seed, 1 if opaque else 0, out.ctypes.data)
    return out


def expand_palette(idx: np.ndarray, palette: np.ndarray) -> np.ndarray:
    """u8 index volume -> RGBA8 texels (uint32 per voxel, r in the low byte)."""
    idx = np.ascontiguousarray(idx)
    out = np.empty(idx.shape, dtype=np.uint32)
    scene_lib().vt_expand_palette(idx.ctypes.data, idx.size, palette.ctypes.data, out.ctypes.data)
    return out


@dataclass
class Volume:
    data: np.ndarray              # (dz, dy, dx) uint8 index or uint32 RGBA8
    palette: np.ndarray | None    # 256 x uint32 for index volumes

    @property
    def dims(self):
        dz, dy, dx = self.data.shape
        return (dx, dy, dz)

    @property
    def fmt(self) -> int:
        return 0 if self.data.dtype == np.uint8 else 1


@dataclass
class RenderParams:
    width: int
    height: int
    viewport_square: int = 1       # the reference's W x W viewport (lib/command.c:77-84)
    mode: int = 0                  # 0 primary colour, 1 primary + shadow ray, 2 path tracing
    spp: int = 1
    max_bounces: int = 0
    jitter: int = 0
    seed: int = 0
    frame: int = 0
    sun_dir: tuple = (0.0, 1.0, 0.0)
    sun_radiance: tuple = (2.5, 2.5, 2.5)
    sky: tuple = SKY


@dataclass
class Scene:
    volumes: list
    records: np.ndarray           # (n, 16) float32 instance records
    P: np.ndarray
    C: np.ndarray
    params: RenderParams
    name: str = ""
    info: dict = field(default_factory=dict)


def sun_direction(rng: Rng) -> tuple:
    """Seeded unit vector towards the sun, elevation 30-70 degrees above the y-down 'up'.

    The reference world is y-down (a GL projection used unflipped in Vulkan,
    src/world.rs:120-125 moves 'up' along -y), so the sun sits at negative y."""
    az = rng.uniform(0.0, 2.0 * math.pi)
    el = rng.uniform(math.radians(30.0), math.radians(70.0))
    d = np.array([math.cos(el) * math.cos(az), -math.sin(el), math.cos(el) * math.sin(az)])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    return tuple(float(x) for x in d)


def orbit_camera(rng: Rng, centre, radius: float, jitter: float):
    theta = rng.uniform(0.0, 2.0 * math.pi)
    phi = rng.uniform(math.radians(35.0), math.radians(145.0))
    eye = np.array(centre, dtype=np.float64) + radius * np.array(
        [math.sin(phi) * math.cos(theta), math.cos(phi), math.sin(phi) * math.sin(theta)])
    target = np.array(centre, dtype=np.float64) + np.array(
        [rng.uniform(-jitter, jitter) for _ in range(3)])
    return eye, target


def single_volume_scene(n: int = 128, fill: float = 0.30, seed: int = 1, width: int = 640, height: int = 480,
                        mode: int = 1, spp: int = 1, max_bounces: int = 0, jitter: int = 0,
                        viewport_square: int = 1, opaque: bool = True, name: str = "") -> Scene:
    """One n^3 noise volume at the origin, camera on a sphere of radius 1.5 n outside it."""
    rng = Rng(seed)
    vol = Volume(gen_volume((n, n, n), seed, fill), gen_palette(seed, opaque))
    rec = instance_record(model_matrix((0.0, 0.0, 0.0), float(n)), 0)[None, :]
    eye, target = orbit_camera(rng, (0.0, 0.0, 0.0), 1.5 * n, 0.05 * n)
    P = perspective(FOV, width / height)
    C = look_at(eye, target)
    prm = RenderParams(width, height, viewport_square=viewport_square, mode=mode, spp=spp,
                       max_bounces=max_bounces, jitter=jitter, seed=seed, sun_dir=sun_direction(rng))
    return Scene([vol], rec, P, C, prm, name=name or f"single{n}", info={"eye": eye, "target": target})


def config1(width: int = 640, height: int = 480) -> Scene:
    """128^3, 30% fill, 640x480, 1 spp at pixel centres, primary + 1 shadow ray."""
    return single_volume_scene(128, 0.30, 1, width, height, mode=1, name="c1_128cube_640x480_1spp_shadow")


def _place_boxes(rng: Rng, count: int, size: int, world: int, margin: int = 8):
    boxes = []
    tries = 0
    while len(boxes) < count:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place boxes without overlap")
        t = np.array([rng.randint(0, world - size) for _ in range(3)], dtype=np.float64)
        ok = all(np.any(np.abs(t - b) >= size + margin) for b in boxes)
        if ok:
            boxes.append(t)
    return boxes


def multi_volume_scene(count: int = 16, n: int = 256, world: int = 2048, seed: int = 2, width: int = 1920,
                       height: int = 1080, spp: int = 4, max_bounces: int = 4, fill_range=(0.10, 0.40),
                       viewport_square: int = 0, name: str = "") -> Scene:
    """count noise volumes of n^3 (fill U[fill_range]) at non-overlapping random places in a
    world^3 cube; camera inside the world looking at a seeded instance."""
    rng = Rng(seed)
    vols, recs = [], []
    corners = _place_boxes(rng, count, n, world)
    for i in range(count):
        fill = rng.uniform(*fill_range)
        vols.append(Volume(gen_volume((n, n, n), seed * 1000 + i, fill), gen_palette(seed * 1000 + i)))
        centre = corners[i] + n / 2.0
        recs.append(instance_record(model_matrix(tuple(centre), float(n)), i))
    centres = np.array([c + n / 2.0 for c in corners])
    # eye: 1.1 n from a seeded target instance centre, on the side away from the cluster centre,
    # looking through the target towards the rest of the world (clear of every box)
    mid = centres.mean(axis=0)
    order = [rng.randint(0, count - 1)] + list(range(count))
    eye, target = None, None
    for ti in order:
        out = centres[ti] - mid
        if np.linalg.norm(out) < 1e-6:
            out = np.array([1.0, -0.5, 0.3])
        out /= np.linalg.norm(out)
        jit = np.array([rng.uniform(-0.25, 0.25) for _ in range(3)])
        dirv = out + jit
        dirv /= np.linalg.norm(dirv)
        cand = centres[ti] + 1.1 * n * dirv
        if all(np.max(np.abs(cand - c)) > 0.75 * n for c in centres):
            eye, target = cand, centres[ti] - 0.25 * n * dirv
            break
    if eye is None:
        raise RuntimeError("no clear camera position")
    P = perspective(FOV, width / height)
    C = look_at(eye, target)
    prm = RenderParams(width, height, viewport_square=viewport_square, mode=2, spp=spp, max_bounces=max_bounces,
                       jitter=1, seed=seed, sun_dir=sun_direction(rng))
    return Scene(vols, np.stack(recs), P, C, prm,
                 name=name or f"c2_{count}x{n}cube_{width}x{height}_{spp}spp_{max_bounces}b",
                 info={"eye": eye, "target": target})


def config2(width: int = 1920, height: int = 1080, spp: int = 4) -> Scene:
    """16 x 256^3 (fill U[10%,40%]) in a 2048^3 world, 1920x1080, 4 spp, <= 4 diffuse bounces."""
    return multi_volume_scene(16, 256, 2048, 2, width, height, spp, 4)


def config3(width: int = 3840, height: int = 2160, n: int = 1024) -> Scene:
    """Single n^3 volume at 5% fill (noise caves + 64 seeded boxes), 1 spp, 2 bounces."""
    rng = Rng(3)
    boxes = []
    for _ in range(64):
        s = [rng.randint(max(2, n // 128), max(3, n // 16)) for _ in range(3)]
        o = [rng.randint(0, n - s[a]) for a in range(3)]
        boxes.append([o[0], o[1], o[2], o[0] + s[0], o[1] + s[1], o[2] + s[2]])
    vol = Volume(gen_volume((n, n, n), 3, 0.05, boxes=boxes), gen_palette(3))
    rec = instance_record(model_matrix((0.0, 0.0, 0.0), float(n)), 0)[None, :]
    eye, target = orbit_camera(rng, (0.0, 0.0, 0.0), 1.2 * n, 0.05 * n)
    P = perspective(FOV, width / height)
    C = look_at(eye, target)
    prm = RenderParams(width, height, viewport_square=0, mode=2, spp=1, max_bounces=2, jitter=1, seed=3,
                       sun_dir=sun_direction(rng))
    return Scene([vol], rec, P, C, prm, name=f"c3_{n}cube_5pct_{width}x{height}_1spp_2b",
                 info={"eye": eye, "target": target})
