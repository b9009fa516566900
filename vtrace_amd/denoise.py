"""Learned denoiser: thin wrapper over the vt_dn_* functions of librender_hip (denoise.hip).

The network (5-level conv U-Net, BF16 MFMA with FP32 accumulation), its training loop
(Adam on random crops of device-resident frames) and inference all run in the library on
the GPU; this module declares the C functions, initialises parameters, renders training
pairs with the plain path and measures quality (relMSE against a high-spp render).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import scenes
from .renderer import RendererError

ENC = (32, 32, 48, 64, 96)
DEC = (64, 48, 32, 32)
N_FEATURES = 11   # of the 16 input channels; the rest are zero padding


class VtDnConfig(C.Structure):
    _fields_ = [("enc", C.c_uint32 * 5), ("dec", C.c_uint32 * 4)]


def declare(lib):
    if getattr(lib, "_vt_dn_declared", False):
        return
    f = C.POINTER(C.c_float)
    lib.vt_dn_init.restype = C.c_int32
    lib.vt_dn_init.argtypes = [C.POINTER(VtDnConfig)]
    lib.vt_dn_layers.restype = C.c_int32
    lib.vt_dn_layers.argtypes = [C.c_void_p, C.c_uint32]
    lib.vt_dn_param_count.restype = C.c_uint64
    lib.vt_dn_param_count.argtypes = []
    lib.vt_dn_set_params.restype = C.c_int32
    lib.vt_dn_set_params.argtypes = [C.c_void_p, C.c_uint64]
    lib.vt_dn_get_params.restype = C.c_int32
    lib.vt_dn_get_params.argtypes = [C.c_void_p, C.c_uint64]
    lib.vt_dn_grad.restype = C.c_int32
    lib.vt_dn_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, f]
    lib.vt_dn_features.restype = C.c_int32
    lib.vt_dn_features.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.vt_dn_add_frame.restype = C.c_int32
    lib.vt_dn_add_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.vt_dn_clear_frames.restype = C.c_int32
    lib.vt_dn_clear_frames.argtypes = []
    lib.vt_dn_train.restype = C.c_int32
    lib.vt_dn_train.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_void_p]
    lib.vt_dn_denoise.restype = C.c_int32
    lib.vt_dn_denoise.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, f]
    lib.vt_dn_denoise_frame.restype = C.c_int32
    lib.vt_dn_denoise_frame.argtypes = [C.c_void_p, f]
    lib.vt_dn_set_blend.restype = C.c_int32
    lib.vt_dn_set_blend.argtypes = [C.c_void_p, C.c_uint32]
    lib.vt_dn_shutdown.restype = None
    lib.vt_dn_shutdown.argtypes = []
    lib._vt_dn_declared = True


def relmse(x: np.ndarray, ref: np.ndarray) -> float:
    """mean((x - r)^2 / (r^2 + 0.01)) over pixels and RGB."""
    x = np.asarray(x, dtype=np.float64)[..., :3]
    r = np.asarray(ref, dtype=np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + 0.01)))


def blend_weight(noisy: np.ndarray, den: np.ndarray, ref: np.ndarray) -> float:
    """alpha in [0, 1] minimising mean((alpha d + (1 - alpha) n - r)^2 / (r^2 + 0.01)) (the relMSE
    of the blend): the zero of its derivative, sum w (n - r)(n - d) / sum w (n - d)^2."""
    n, d, r = (np.asarray(v, dtype=np.float64)[..., :3] for v in (noisy, den, ref))
    w = 1.0 / (r * r + 0.01)
    den_ = float(np.sum(w * (n - d) ** 2))
    if den_ <= 0.0:
        return 1.0
    return float(np.clip(np.sum(w * (n - r) * (n - d)) / den_, 0.0, 1.0))


class Denoiser:
    def __init__(self, renderer, enc=ENC, dec=DEC, seed: int | None = 0):
        self.r = renderer
        self.lib = renderer.lib
        declare(self.lib)
        self.enc, self.dec = tuple(enc), tuple(dec)
        cfg = VtDnConfig((C.c_uint32 * 5)(*self.enc), (C.c_uint32 * 4)(*self.dec))
        if self.lib.vt_dn_init(C.byref(cfg)) != 0:
            raise RendererError(f"vt_dn_init failed: {renderer.error()}")
        tab = np.zeros((32, 4), dtype=np.uint32)
        n = self.lib.vt_dn_layers(tab.ctypes.data, 32)
        self.layers = [tuple(int(v) for v in tab[i]) for i in range(n)]
        self.nparam = int(self.lib.vt_dn_param_count())
        if seed is not None:
            self.set_params(self.init_params(seed))

    @property
    def real_param_count(self) -> int:
        """Parameters excluding the zero padding of the input channels."""
        cin, cout = self.layers[0][0], self.layers[0][1]
        return self.nparam - cout * 9 * (cin - N_FEATURES)

    def init_params(self, seed: int) -> np.ndarray:
        """He-normal weights (fan-in of the real inputs), zero biases, zero padded channels."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparam, dtype=np.float32)
        for i, (cin, cout, wo, bo) in enumerate(self.layers):
            fan = 9 * (N_FEATURES if i == 0 else cin)
            std = math.sqrt(2.0 / fan) if i < len(self.layers) - 1 else math.sqrt(1.0 / fan)
            w = rng.normal(0.0, std, size=(cout, 3, 3, cin)).astype(np.float32)
            if i == 0:
                w[..., N_FEATURES:] = 0.0
            p[wo:wo + w.size] = w.reshape(-1)
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam or self.lib.vt_dn_set_params(p.ctypes.data, p.size) != 0:
            raise RendererError("vt_dn_set_params failed")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparam, dtype=np.float32)
        if self.lib.vt_dn_get_params(p.ctypes.data, p.size) != 0:
            raise RendererError("vt_dn_get_params failed")
        return p

    def grad(self, x: np.ndarray, target: np.ndarray):
        """x (n,h,w,16), target (n,h,w,3) -> (y, grads, loss) of one forward+backward pass."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        n, h, w, _ = x.shape
        y = np.zeros((n, h, w, 3), dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float(0.0)
        if self.lib.vt_dn_grad(x.ctypes.data, t.ctypes.data, n, h, w, y.ctypes.data, g.ctypes.data,
                               C.byref(loss)) != 0:
            raise RendererError(f"vt_dn_grad failed: {self.r.error()}")
        return y, g, float(loss.value)

    def features(self, rad: np.ndarray, gbuf: np.ndarray, spp: int = 1) -> np.ndarray:
        rad = np.ascontiguousarray(rad, dtype=np.float32)
        gbuf = np.ascontiguousarray(gbuf, dtype=np.float32)
        h, w = rad.shape[:2]
        out = np.zeros((h, w, 16), dtype=np.float32)
        if self.lib.vt_dn_features(rad.ctypes.data, gbuf.ctypes.data, h, w, spp, out.ctypes.data) != 0:
            raise RendererError("vt_dn_features failed")
        return out

    def add_frame(self, rad: np.ndarray, gbuf: np.ndarray, target: np.ndarray, spp: int) -> int:
        rad = np.ascontiguousarray(rad, dtype=np.float32)
        gbuf = np.ascontiguousarray(gbuf, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        h, w = rad.shape[:2]
        k = self.lib.vt_dn_add_frame(rad.ctypes.data, gbuf.ctypes.data, target.ctypes.data, h, w, spp)
        if k < 0:
            raise RendererError(f"vt_dn_add_frame failed: {self.r.error()}")
        return k

    def clear_frames(self):
        self.lib.vt_dn_clear_frames()

    def train(self, iters: int, batch: int = 8, crop: int = 128, lr: float = 2e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(iters, dtype=np.float32)
        if self.lib.vt_dn_train(iters, batch, crop, lr, seed, losses.ctypes.data) != 0:
            raise RendererError(f"vt_dn_train failed: {self.r.error()}")
        return losses

    def denoise(self, rad: np.ndarray, gbuf: np.ndarray, spp: int):
        rad = np.ascontiguousarray(rad, dtype=np.float32)
        gbuf = np.ascontiguousarray(gbuf, dtype=np.float32)
        h, w = rad.shape[:2]
        out = np.zeros((h, w, 4), dtype=np.float32)
        ms = C.c_float(0.0)
        if self.lib.vt_dn_denoise(rad.ctypes.data, gbuf.ctypes.data, h, w, spp, out.ctypes.data, C.byref(ms)) != 0:
            raise RendererError(f"vt_dn_denoise failed: {self.r.error()}")
        return out, float(ms.value)

    def set_blend(self, alphas) -> None:
        """Output blend weights for spp = 1, 2, 4, ... (vt_dn_set_blend); [] removes them."""
        a = np.ascontiguousarray(np.asarray(alphas, dtype=np.float32).reshape(-1))
        if self.lib.vt_dn_set_blend(a.ctypes.data if a.size else None, a.size) != 0:
            raise RendererError(f"vt_dn_set_blend failed: {self.r.error()}")

    def calibrate_blend(self, r, scene: "scenes.Scene", max_log2_spp: int = 8, target_spp: int = 1024,
                        seed: int = 31) -> list:
        """Per spp = 2^k (k <= max_log2_spp), the blend weight that minimises the relMSE of
        alpha * network output + (1 - alpha) * input on `scene` against a target_spp render of it
        (a scene kept out of training and evaluation); sets and returns the table."""
        from .renderer import load_scene
        load_scene(r, scene)
        target, _, _ = render_spp(r, scene, target_spp, seed=seed * 1000 + 7)
        self.set_blend([])
        alphas = []
        for k in range(max_log2_spp + 1):
            spp = 1 << k
            rad, gb, _ = render_spp(r, scene, spp, seed=seed * 1000 + 100 + k, gbuffer=True)
            out, _ = self.denoise(rad, gb, spp)
            alphas.append(blend_weight(rad, out, target))
        self.set_blend(alphas)
        return alphas

    def denoise_frame(self):
        """Denoises the renderer's last frame (G-buffer enabled) without a host round trip."""
        w, h = self.r.window_width, self.r.window_height
        out = np.zeros((h, w, 4), dtype=np.float32)
        ms = C.c_float(0.0)
        if self.lib.vt_dn_denoise_frame(out.ctypes.data, C.byref(ms)) != 0:
            raise RendererError(f"vt_dn_denoise_frame failed: {self.r.error()}")
        return out, float(ms.value)

    def close(self):
        self.lib.vt_dn_shutdown()


# ---- rendering helpers ----------------------------------------------------------------------
def render_spp(r, scene: "scenes.Scene", spp: int, seed: int, chunk: int = 16, gbuffer: bool = False):
    """Radiance of `spp` samples per pixel as the mean of frames of `chunk` samples (frame
    index = independent RNG stream).  Returns (radiance (h,w,4), G-buffer or None, GPU ms)."""
    p = scene.params
    per = min(spp, chunk)
    nfr = max(1, spp // per)
    r.enable_gbuffer(gbuffer)
    acc = None
    gb = None
    ms = 0.0
    for k in range(nfr):
        prm = scenes.RenderParams(p.width, p.height, viewport_square=p.viewport_square, mode=2, spp=per,
                                  max_bounces=p.max_bounces, jitter=1, seed=seed, frame=k, sun_dir=p.sun_dir,
                                  sun_radiance=p.sun_radiance, sky=p.sky)
        r.set_params(prm)
        r.render(scene.P, scene.C)
        ms += r.last_frame_ms()
        rad = r.radiance().astype(np.float64)
        acc = rad if acc is None else acc + rad
        if gbuffer and k == 0:
            gb = r.gbuffer()
    r.enable_gbuffer(False)
    return (acc / nfr).astype(np.float32), gb, ms


def crossing_at(curve: list, threshold: float):
    """Time (ms) at which a curve of {"ms", "relmse"} points, relMSE falling, reaches threshold
    (log-log interpolation; None if it never does)."""
    prev = None
    for c in curve:
        if c["relmse"] <= threshold:
            if prev is None:
                return c["ms"]
            f = (math.log(prev["relmse"]) - math.log(threshold)) / (math.log(prev["relmse"]) - math.log(c["relmse"]))
            return math.exp(math.log(prev["ms"]) + f * (math.log(c["ms"]) - math.log(prev["ms"])))
        prev = c
    return None


# ---- benchmark ------------------------------------------------------------------------------
def benchmark(r, eval_scene: "scenes.Scene", train_seeds=(101, 102, 103, 104, 105, 106),
              train_spp=(1, 4, 16, 64, 256), target_spp: int = 1024, judge_spp: int = 4096, iters: int = 24000,
              batch: int = 8, crop: int = 128,
              lr: float = 2e-3, threshold: float = 1e-3, calib_seed: int = 107, log=None) -> dict:
    """Trains on C2-style scenes (other seeds than the evaluation scene), calibrates the output
    blend on one more seed, then measures the GPU time each method needs to reach relMSE <=
    threshold against a judge_spp render of eval_scene: the plain path (spp doubling from 1,
    frames without G-buffer) and render + denoise, network alone and blended (spp doubling
    from 1).  Times are device times (render kernels + denoiser kernels)."""
    say = log or (lambda *a: None)
    p = eval_scene.params
    d = Denoiser(r, seed=0)
    t0 = time.perf_counter()
    for s in train_seeds:
        sc = scenes.multi_volume_scene(16, 256, 2048, s, p.width, p.height, 1, p.max_bounces)
        from .renderer import load_scene
        load_scene(r, sc)
        del sc.volumes[:]
        target, _, _ = render_spp(r, sc, target_spp, seed=90000 + s)
        for spp in train_spp:
            inp, gb, _ = render_spp(r, sc, spp, seed=1000 * s + spp, gbuffer=True)
            d.add_frame(inp, gb, target, spp)
    t_data = time.perf_counter() - t0
    say(f"denoiser: {len(train_seeds) * len(train_spp)} training frames in {t_data:.1f} s")
    t0 = time.perf_counter()
    losses = d.train(iters, batch=batch, crop=crop, lr=lr, seed=7)
    t_train = time.perf_counter() - t0
    d.clear_frames()
    say(f"denoiser: {iters} iterations in {t_train:.1f} s, loss {losses[:50].mean():.4f} -> {losses[-50:].mean():.4f}")
    # output blend weights, calibrated on a scene of its own (neither trained on nor evaluated)
    calib = scenes.multi_volume_scene(16, 256, 2048, calib_seed, p.width, p.height, 1, p.max_bounces)
    alphas = d.calibrate_blend(r, calib, max_log2_spp=8, target_spp=target_spp, seed=calib_seed)
    del calib.volumes[:]
    say("denoiser: blend weights " + ", ".join(f"{1 << k}:{a:.2f}" for k, a in enumerate(alphas)))
    from .renderer import load_scene
    load_scene(r, eval_scene)
    judge, _, judge_ms = render_spp(r, eval_scene, judge_spp, seed=777)
    plain, den, blended = [], [], []
    spp = 1
    while spp <= target_spp:
        # plain: the frame alone (no G-buffer); the denoiser's input: the same samples with it
        rad_plain, _, ms_plain = render_spp(r, eval_scene, spp, seed=4242)
        e = relmse(rad_plain, judge)
        plain.append({"spp": spp, "ms": round(ms_plain, 3), "relmse": e})
        if spp <= 256:
            rad, gb, ms = render_spp(r, eval_scene, spp, seed=4242, gbuffer=True)
            d.set_blend([])
            out, dms = d.denoise(rad, gb, spp)
            den.append({"spp": spp, "ms": round(ms + dms, 3), "render_ms": round(ms, 3), "denoise_ms": round(dms, 4),
                        "relmse": relmse(out, judge)})
            d.set_blend(alphas)
            outb, bms = d.denoise(rad, gb, spp)
            blended.append({"spp": spp, "ms": round(ms + bms, 3), "render_ms": round(ms, 3),
                            "denoise_ms": round(bms, 4), "relmse": relmse(outb, judge)})
        say(f"denoiser: spp {spp}: plain {e:.3e}" +
            (f", denoised {den[-1]['relmse']:.3e}, blended {blended[-1]['relmse']:.3e}" if spp <= 256 else ""))
        if e <= threshold and den and den[-1]["relmse"] <= threshold and blended[-1]["relmse"] <= threshold:
            break
        spp *= 2

    def crossing(curve):
        """(ms, spp) where relMSE reaches threshold: log-log interpolation between the last
        point above and the first point at or below it (the first point if it is already below)."""
        prev = None
        for c in curve:
            if c["relmse"] <= threshold:
                if prev is None:
                    return c["ms"], c["spp"], False
                f = (math.log(prev["relmse"]) - math.log(threshold)) / (math.log(prev["relmse"]) - math.log(c["relmse"]))
                ms = math.exp(math.log(prev["ms"]) + f * (math.log(c["ms"]) - math.log(prev["ms"])))
                spp = math.exp(math.log(prev["spp"]) + f * (math.log(c["spp"]) - math.log(prev["spp"])))
                return ms, spp, False
            prev = c
        if len(curve) >= 2:  # Monte Carlo error falls as 1/spp: extrapolate from the last point
            last = curve[-1]
            k = last["relmse"] / threshold
            return last["ms"] * k, last["spp"] * k, True
        return None

    out = {"judge_spp": judge_spp, "judge_ms": round(judge_ms, 1), "threshold": threshold,
           "params": d.real_param_count, "train": {"frames": len(train_seeds) * len(train_spp), "iters": iters,
                                                   "batch": batch, "crop": crop, "seconds": round(t_train, 2),
                                                   "data_seconds": round(t_data, 1),
                                                   "loss_first50": float(losses[:50].mean()),
                                                   "loss_last200": float(losses[-200:].mean())},
           "blend_weights": [round(a, 4) for a in alphas], "blend_calibration_seed": calib_seed,
           "plain_curve": plain, "denoised_curve": den, "blended_curve": blended}
    for name, curve in (("plain", plain), ("denoised", den), ("blended", blended)):
        c = crossing(curve)
        out[f"{name}_time_to_threshold_ms"] = round(c[0], 3) if c else None
        out[f"{name}_spp_at_threshold"] = round(c[1], 1) if c else None
        out[f"{name}_extrapolated"] = bool(c[2]) if c else None
    out["denoised_best_relmse"] = min(c["relmse"] for c in den) if den else None
    if out["plain_time_to_threshold_ms"] and out["denoised_time_to_threshold_ms"]:
        out["speedup_network_alone"] = round(out["plain_time_to_threshold_ms"] / out["denoised_time_to_threshold_ms"], 3)
    if out["plain_time_to_threshold_ms"] and out["blended_time_to_threshold_ms"]:
        out["speedup"] = round(out["plain_time_to_threshold_ms"] / out["blended_time_to_threshold_ms"], 3)
    # equal-quality speedups at low spp: plain time to reach the denoised error of each input spp
    eq = []
    for c in den:
        sub = [q for q in plain if q["relmse"] > 0]
        tgt = c["relmse"]
        t_plain = crossing_at(sub, tgt)
        if t_plain:
            eq.append({"spp": c["spp"], "denoised_ms": c["ms"], "plain_ms_same_relmse": round(t_plain, 3)})
    out["equal_quality"] = eq
    d.close()
    return out
