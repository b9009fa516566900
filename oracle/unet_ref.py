"""FP64 CPU reference of the denoiser network (forward, relMSE loss, backward).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench-side checks as the checker of the
HIP denoiser (vtrace_amd/csrc/denoise.hip); the product never imports it.  The denoiser has
no counterpart in the reference repository, so this module follows the network's own
specification, restated here:

* tensors are NHWC; a 3x3 convolution (zero padding 1) of layer i has weights
  W[co][ky][kx][ci] and bias b[co] stored in one flat parameter vector at the offsets of the
  library's layer table (vt_dn_layers);
* encoder level l (l = 0..3): conv(2l) + ReLU, conv(2l+1) + ReLU (= skip_l), 2x2 mean pooling;
  bottleneck: conv(8) + ReLU, conv(9) + ReLU;
* decoder level l (l = 3..0): nearest 2x upsampling of the coarser output, concatenation
  [upsampled, skip_l] along channels, conv + ReLU, conv + ReLU; output conv(18), linear;
* the network predicts a correction y to the input's log irradiance b (input channels 0..2):
  irradiance E = exp(min(b + y, 16)) - 1, radiance a * E with a = albedo + 0.01 (albedo =
  input channels 3..5); loss = mean over pixels and RGB of (a E - t)^2 / (t^2 + 0.01).

`bf16=True` rounds to BF16 (round to nearest even) at the points where the GPU stores BF16
tensors: the input, every stored activation and activation gradient, the weights; the
output, biases, loss and weight gradients stay unrounded, as on the GPU.  `bf16=False` is the
plain FP64 reference.  The FP64 mode is pinned by finite differences and by an independent
convolution (scipy.signal.correlate) in tests/test_denoise_ref.py.
"""
from __future__ import annotations

import numpy as np

ALBEDO_EPS = 0.01
LOG_MAX = 16.0


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _im2col(x: np.ndarray) -> np.ndarray:
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c))
    xp[:, 1:-1, 1:-1, :] = x
    cols = np.empty((n, h, w, 9, c))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx, :] = xp[:, ky:ky + h, kx:kx + w, :]
    return cols.reshape(n * h * w, 9 * c)


def _mm(a, b, acc32):
    if acc32:
        return (a.astype(np.float32) @ b.astype(np.float32)).astype(np.float64)
    return a @ b


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray, acc32: bool = False) -> np.ndarray:
    """x (n,h,w,ci), w (co,3,3,ci), b (co) -> (n,h,w,co).  acc32: FP32 accumulation."""
    n, h, ww, _ = x.shape
    co = w.shape[0]
    y = _mm(_im2col(x), w.reshape(co, -1).T, acc32)
    if acc32:
        return (y.astype(np.float32) + b.astype(np.float32)).astype(np.float64).reshape(n, h, ww, co)
    return (y + b).reshape(n, h, ww, co)


def conv3x3_backward(x: np.ndarray, w: np.ndarray, g: np.ndarray, acc32: bool = False):
    """Gradients of conv3x3 for output gradient g: (dx, dw, db)."""
    n, h, ww, ci = x.shape
    co = w.shape[0]
    g2 = g.reshape(-1, co)
    dw = _mm(g2.T, _im2col(x), acc32).reshape(co, 3, 3, ci)
    db = g2.sum(axis=0)
    dcols = _mm(g2, w.reshape(co, -1), acc32).reshape(n, h, ww, 9, ci)
    dxp = np.zeros((n, h + 2, ww + 2, ci))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + h, kx:kx + ww, :] += dcols[:, :, :, ky * 3 + kx, :]
    return dxp[:, 1:-1, 1:-1, :], dw, db


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool2_backward(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_backward(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def layer_table(enc, dec, in_ch: int = 16, out_ch: int = 3):
    """(cin, cout, w_off, b_off) of the 19 layers in parameter-vector order."""
    E, D = list(enc), list(dec)
    U = [D[1], D[2], D[3], E[4]]
    shapes = [(in_ch, E[0]), (E[0], E[0])]
    for l in range(1, 4):
        shapes += [(E[l - 1], E[l]), (E[l], E[l])]
    shapes += [(E[3], E[4]), (E[4], E[4])]
    for l in (3, 2, 1, 0):
        shapes += [(U[l] + E[l], D[l]), (D[l], D[l])]
    shapes.append((D[0], out_ch))
    out, off = [], 0
    for cin, cout in shapes:
        out.append((cin, cout, off, off + cout * 9 * cin))
        off += cout * 9 * cin + cout
    return out


class UNetRef:
    """layers: list of (cin, cout, w_off, b_off) as returned by the library's layer table."""

    def __init__(self, layers, enc, dec, acc32: bool = False):
        """acc32: accumulate convolutions in FP32 instead of FP64 (a second, equally valid
        summation; the spread between the two bounds what summation order alone can change
        once BF16 rounding is in the loop)."""
        self.acc32 = acc32
        self.layers = [tuple(int(v) for v in l) for l in layers]
        self.E = list(enc)
        self.D = list(dec)
        self.U = [self.D[1], self.D[2], self.D[3], self.E[4]]

    def weights(self, params, i, bf16):
        cin, cout, wo, bo = self.layers[i]
        w = np.asarray(params[wo:wo + cout * 9 * cin], dtype=np.float64).reshape(cout, 3, 3, cin)
        b = np.asarray(params[bo:bo + cout], dtype=np.float64)
        return (bf16_round(w) if bf16 else w), b

    def forward(self, params, x, bf16=False):
        R = bf16_round if bf16 else (lambda a: a)
        c = {"x0": R(x)}
        relu = lambda a: np.maximum(a, 0.0)
        h = c["x0"]
        for l in range(4):
            w, b = self.weights(params, 2 * l, bf16)
            c[f"a{l}"] = R(relu(conv3x3(h, w, b, self.acc32)))
            w, b = self.weights(params, 2 * l + 1, bf16)
            c[f"s{l}"] = R(relu(conv3x3(c[f"a{l}"], w, b, self.acc32)))
            c[f"p{l}"] = R(pool2(c[f"s{l}"]))
            h = c[f"p{l}"]
        w, b = self.weights(params, 8, bf16)
        c["b1"] = R(relu(conv3x3(h, w, b, self.acc32)))
        w, b = self.weights(params, 9, bf16)
        c["b2"] = R(relu(conv3x3(c["b1"], w, b, self.acc32)))
        prev = c["b2"]
        for l in (3, 2, 1, 0):
            li = 10 + 2 * (3 - l)
            c[f"cat{l}"] = np.concatenate([up2(prev), c[f"s{l}"]], axis=3)
            w, b = self.weights(params, li, bf16)
            c[f"d1{l}"] = R(relu(conv3x3(c[f"cat{l}"], w, b, self.acc32)))
            w, b = self.weights(params, li + 1, bf16)
            c[f"d2{l}"] = R(relu(conv3x3(c[f"d1{l}"], w, b, self.acc32)))
            prev = c[f"d2{l}"]
        w, b = self.weights(params, 18, bf16)
        y = conv3x3(prev, w, b, self.acc32)
        return y, c

    @staticmethod
    def loss(y, t, x):
        """x: the (rounded) input features; returns (loss, dLoss/dy)."""
        b, a = x[..., 0:3], x[..., 3:6] + ALBEDO_EPS
        z = b + y
        dE = np.where(z > LOG_MAX, 0.0, np.exp(np.minimum(z, LOG_MAX)))
        E = np.expm1(np.minimum(z, LOG_MAX))
        den = t * t + 0.01
        d = a * E - t
        return float(np.mean(d * d / den)), 2.0 * d / den * a * dE / d.size

    def backward(self, params, c, gy, bf16=False):
        R = bf16_round if bf16 else (lambda a: a)
        grads = np.zeros(len(params))

        def put(i, dw, db):
            cin, cout, wo, bo = self.layers[i]
            grads[wo:wo + cout * 9 * cin] = dw.reshape(-1)
            grads[bo:bo + cout] = db

        def conv_back(i, x, g, mask=None, need_dx=True):
            w, _ = self.weights(params, i, bf16)
            dx, dw, db = conv3x3_backward(x, w, g, self.acc32)
            put(i, dw, db)
            if not need_dx:
                return None
            if mask is not None:
                dx = dx * (mask > 0)
            return R(dx)

        g = R(gy)
        g = conv_back(18, c["d20"], g, c["d20"])
        gcat = {}
        for l in (0, 1, 2, 3):
            li = 10 + 2 * (3 - l)
            g1 = conv_back(li + 1, c[f"d1{l}"], g, c[f"d1{l}"])
            gc = conv_back(li, c[f"cat{l}"], g1, c[f"cat{l}"])
            gcat[l] = gc
            g = R(up2_backward(gc[..., :self.U[l]]))
        g = conv_back(9, c["b1"], g, c["b1"])
        g = conv_back(8, c["p3"], g, None)
        for l in (3, 2, 1, 0):
            gs = R((pool2_backward(g) + gcat[l][..., self.U[l]:]) * (c[f"s{l}"] > 0))
            ga = conv_back(2 * l + 1, c[f"a{l}"], gs, c[f"a{l}"])
            x_in = c[f"p{l - 1}"] if l > 0 else c["x0"]
            g = conv_back(2 * l, x_in, ga, None, need_dx=l > 0)
        return grads

    def loss_and_grads(self, params, x, t, bf16=False):
        y, c = self.forward(params, x, bf16)
        loss, gy = self.loss(y, t, c["x0"])
        return y, loss, self.backward(params, c, gy, bf16)


def features(rad_rgba: np.ndarray, gbuf8: np.ndarray, spp: int = 1) -> np.ndarray:
    """The 16 input channels of a frame of spp samples (FP64, before BF16 rounding):
    demodulated log irradiance (with the BF16-rounded albedo the network remodulates with),
    albedo, normal, log distance, log2(spp)."""
    h, w = rad_rgba.shape[:2]
    f = np.zeros((h, w, 16))
    a = bf16_round(gbuf8[..., 0:3]) + ALBEDO_EPS
    f[..., 0:3] = np.log1p(np.maximum(rad_rgba[..., :3].astype(np.float64), 0.0) / a)
    f[..., 3:6] = gbuf8[..., 0:3]
    f[..., 6:9] = gbuf8[..., 4:7]
    f[..., 9] = np.log1p(np.maximum(gbuf8[..., 3].astype(np.float64), 0.0)) * 0.125
    f[..., 10] = np.log2(max(spp, 1)) * 0.125
    return f


def radiance(y: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Network output y (log-irradiance correction) -> radiance, remodulated with the albedo
    of the (rounded) features x."""
    E = np.expm1(np.minimum(x[..., 0:3] + y, LOG_MAX))
    return np.maximum((x[..., 3:6] + ALBEDO_EPS) * E, 0.0)
