"""CPU: pins of the FP64 reference of the denoiser network (oracle/unet_ref.py).

The reference is the checker of the HIP denoiser; here it is checked itself: convolution
against scipy.signal.correlate (an independent implementation), every gradient against
central finite differences in FP64, and BF16 rounding against hand-computed values."""
import os
import sys

import numpy as np
import pytest
from scipy import signal

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402


def test_conv3x3_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 7, 9, 5))
    w = rng.normal(size=(4, 3, 3, 5))
    b = rng.normal(size=4)
    y = U.conv3x3(x, w, b)
    for n in range(2):
        for co in range(4):
            ref = b[co] + sum(signal.correlate(x[n, :, :, ci], w[co, :, :, ci], mode="same") for ci in range(5))
            assert np.allclose(y[n, :, :, co], ref, atol=1e-12)


def test_conv3x3_backward_is_the_adjoint():
    # <conv(x), g> is linear in x and w: its gradients must satisfy the adjoint identities
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 6, 5, 3))
    w = rng.normal(size=(2, 3, 3, 3))
    g = rng.normal(size=(1, 6, 5, 2))
    zb = np.zeros(2)
    dx, dw, db = U.conv3x3_backward(x, w, g)
    assert np.isclose(np.sum(U.conv3x3(x, w, zb) * g), np.sum(dx * x))
    assert np.isclose(np.sum(U.conv3x3(x, w, zb) * g), np.sum(dw * w))
    assert np.allclose(db, g.sum(axis=(0, 1, 2)))


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-39, 1.0 + 2 ** -8 + 2 ** -20])
    want = np.array([1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5, U.bf16_round(np.array([3.0e-39]))[0], 1.0 + 2 ** -7])
    assert np.array_equal(U.bf16_round(v), want)
    # every BF16 value is a fixed point
    r = U.bf16_round(np.random.default_rng(2).normal(size=1000))
    assert np.array_equal(U.bf16_round(r), r)


def _tiny():
    enc, dec = (16, 16, 16, 16, 16), (16, 16, 16, 16)
    layers = U.layer_table(enc, dec)
    n = layers[-1][3] + layers[-1][1]
    rng = np.random.default_rng(3)
    params = np.zeros(n)
    for i, (cin, cout, wo, bo) in enumerate(layers):
        params[wo:wo + cout * 9 * cin] = rng.normal(0.0, np.sqrt(2.0 / (9 * cin)), cout * 9 * cin)
        params[bo:bo + cout] = rng.normal(0.0, 0.05, cout)
    return U.UNetRef(layers, enc, dec), params, rng


def test_layer_table_offsets_are_contiguous():
    layers = U.layer_table((32, 32, 48, 64, 96), (64, 48, 32, 32))
    off = 0
    for cin, cout, wo, bo in layers:
        assert wo == off and bo == wo + cout * 9 * cin
        off = bo + cout
    assert len(layers) == 19 and layers[10][0] == 96 + 64 and layers[16][0] == 48 + 32
    assert 450_000 < off < 600_000  # about 0.5 M parameters


def test_gradients_match_finite_differences():
    net, params, rng = _tiny()
    x = rng.normal(size=(1, 16, 16, 16))
    x[..., 11:] = 0.0
    t = np.abs(rng.normal(size=(1, 16, 16, 3)))
    _, loss, g = net.loss_and_grads(params, x, t)
    # every layer: its first weight, a random weight and its bias
    idx = []
    for cin, cout, wo, bo in net.layers:
        idx += [wo, wo + int(rng.integers(cout * 9 * cin)), bo + int(rng.integers(cout))]
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        l1 = net.loss_and_grads(p1, x, t)[1]
        l2 = net.loss_and_grads(p2, x, t)[1]
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_mode_stays_close_to_fp64():
    # BF16 rounding of the stored tensors perturbs the output by well under a percent; the
    # backward pass's own rounding (on the FP64 forward state) perturbs every layer's gradient
    # by about a percent.  (BF16 in the forward pass also flips a few ReLU masks, which moves
    # the gradients of tiny layers by more: that is a property of the network, not an error.)
    net, params, rng = _tiny()
    x = U.bf16_round(rng.normal(size=(1, 16, 16, 16)))
    t = np.abs(rng.normal(size=(1, 16, 16, 3)))
    y64, c64 = net.forward(params, x)
    y16, _ = net.forward(params, x, bf16=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 3e-2
    _, gy = net.loss(y64, t, c64["x0"])
    g64 = net.backward(params, c64, gy)
    g16 = net.backward(params, c64, gy, bf16=True)
    for cin, cout, wo, bo in net.layers:
        a, b = g64[wo:bo + cout], g16[wo:bo + cout]
        assert np.linalg.norm(b - a) / np.linalg.norm(a) < 3e-2


def test_features_formula():
    rad = np.zeros((2, 3, 4), dtype=np.float32)
    rad[0, 0, :3] = [0.0, 0.5, 1.0]
    rad[0, 1, :3] = [-1.0, 0.5, 2.0]
    gb = np.zeros((2, 3, 8), dtype=np.float32)
    gb[0, 0] = [0.25, 0.5, 0.99, np.e ** 8 - 1.0, 0.0, 1.0, 0.0, 1.0]
    f = U.features(rad, gb, spp=16)
    a = np.array([0.25, 0.5, U.bf16_round(np.array([0.99]))[0]]) + 0.01
    assert np.allclose(f[0, 0, :3], np.log1p(np.array([0.0, 0.5, 1.0]) / a), atol=1e-12)
    assert np.allclose(f[0, 0, 3:11], [0.25, 0.5, 0.99, 0.0, 1.0, 0.0, 1.0, 0.5], atol=1e-6)
    assert f[0, 1, 0] == 0.0 and np.isclose(f[0, 1, 2], np.log(1.0 + 2.0 / 0.01))
    assert np.all(f[..., 11:] == 0.0)
    # a zero correction reproduces the input radiance (remodulation inverts demodulation)
    fb = f.copy()
    fb[..., 3:6] = U.bf16_round(f[..., 3:6])  # the network reads BF16 features
    assert np.allclose(U.radiance(np.zeros((2, 3, 3)), fb)[0, 0], [0.0, 0.5, 1.0], atol=1e-9)


def test_loss_gradient_includes_remodulation():
    rng = np.random.default_rng(4)
    y, t = rng.normal(size=(1, 4, 4, 3)), np.abs(rng.normal(size=(1, 4, 4, 3)))
    x = np.zeros((1, 4, 4, 16))
    x[..., 0:3] = rng.uniform(0.0, 2.0, size=(1, 4, 4, 3))
    x[..., 3:6] = rng.uniform(0.0, 1.0, size=(1, 4, 4, 3))
    loss, gy = U.UNetRef.loss(y, t, x)
    E = np.expm1(x[..., 0:3] + y)
    assert np.isclose(loss, np.mean(((x[..., 3:6] + 0.01) * E - t) ** 2 / (t * t + 0.01)))
    eps = 1e-7
    for idx in [(0, 0, 0, 0), (0, 3, 2, 1), (0, 1, 3, 2)]:
        y1, y2 = y.copy(), y.copy()
        y1[idx] += eps
        y2[idx] -= eps
        fd = (U.UNetRef.loss(y1, t, x)[0] - U.UNetRef.loss(y2, t, x)[0]) / (2 * eps)
        assert abs(fd - gy[idx]) < 1e-6 * max(1.0, abs(fd))


def test_blend_weight_is_the_relmse_minimiser():
    """denoise.blend_weight: the closed form equals the minimum of the blended relMSE over a fine
    grid of weights (clamped to [0, 1]); exact estimates get weight 1."""
    from vtrace_amd import denoise as DN
    rng = np.random.default_rng(3)
    for trial in range(20):
        ref = rng.uniform(0.0, 2.0, size=(24, 32, 4))
        noisy = ref + rng.normal(0.0, 0.3, size=ref.shape)
        den = ref + rng.normal(0.0, rng.uniform(0.05, 0.6), size=ref.shape) + rng.uniform(-0.2, 0.2)
        a = DN.blend_weight(noisy, den, ref)
        grid = np.linspace(0.0, 1.0, 2001)
        errs = [DN.relmse(g * den + (1.0 - g) * noisy, ref) for g in grid]
        best = grid[int(np.argmin(errs))]
        assert abs(a - best) <= 1e-3 + 1e-12, (trial, a, best)
        assert DN.relmse(a * den + (1 - a) * noisy, ref) <= min(errs) + 1e-12
    ref = rng.uniform(0.0, 1.0, size=(8, 8, 4))
    assert DN.blend_weight(ref + 0.1, ref, ref) == 1.0


def test_tile_with_halo_reproduces_the_whole_frame_forward():
    """A tile aligned to the 16-pixel pooling grid with a 144-pixel halo around a block gives
    the block exactly as the whole-frame forward does (the GPU 1080p check relies on this)."""
    enc, dec = (32, 32, 48, 64, 96), (64, 48, 32, 32)
    layers = U.layer_table(enc, dec)
    n = layers[-1][3] + layers[-1][1]
    rng = np.random.default_rng(4)
    p = rng.normal(0.0, 0.05, n)
    net = U.UNetRef(layers, enc, dec)
    x = rng.normal(size=(1, 352, 352, 16))
    y, _ = net.forward(p, x)
    cy = cx = 160
    halo = 144
    t0, t1 = (cy - halo) // 16 * 16, -(-(cy + 32 + halo) // 16) * 16
    yt, _ = net.forward(p, x[:, t0:t1, t0:t1])
    a = yt[0, cy - t0:cy - t0 + 32, cx - t0:cx - t0 + 32]
    b = y[0, cy:cy + 32, cx:cx + 32]
    assert np.allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max())
    # a halo well inside the dependence radius does not
    t0s = (cy - 48) // 16 * 16
    ys, _ = net.forward(p, x[:, t0s:t0s + 128, t0s:t0s + 128])
    assert not np.allclose(ys[0, cy - t0s:cy - t0s + 32, cx - t0s:cx - t0s + 32], b, rtol=0, atol=1e-6)
