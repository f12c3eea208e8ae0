"""Pins the FP64 numpy U-Net (oracle/unet_ref.py), the checker of the warm
start's GPU kernels: analytic gradients against central finite differences,
per parameter and along a random direction through all parameters, and the
layer adjoints (pooling/upsampling, convolution) against their definitions."""
import numpy as np

from oracle import unet_ref as U


def _loss(p, x, t, levels):
    y, _ = U.forward(p, x, levels)
    return U.mse(y, t)[0]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    levels = 3
    p = U.init_params(rng, 2, 2, levels, bf16=False)
    x = rng.standard_normal((2, 8, 8, 2))
    t = rng.standard_normal((2, 8, 8, 1))
    y, cache = U.forward(p, x, levels)
    _, dy = U.mse(y, t)
    g = U.backward(p, cache, dy, levels)
    assert set(g) == set(p)
    h = 1e-6
    checked = 0
    for name, a in p.items():
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in a.shape)
            old = a[idx]
            a[idx] = old + h
            lp = _loss(p, x, t, levels)
            a[idx] = old - h
            lm = _loss(p, x, t, levels)
            a[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (name, idx, fd, g[name][idx])
            checked += fd != 0
    assert checked > 20
    # a random direction through every parameter at once
    v = {k: rng.standard_normal(a.shape) for k, a in p.items()}
    pp = {k: a + h * v[k] for k, a in p.items()}
    pm = {k: a - h * v[k] for k, a in p.items()}
    fd = (_loss(pp, x, t, levels) - _loss(pm, x, t, levels)) / (2 * h)
    an = sum(float((g[k] * v[k]).sum()) for k in p)
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_layer_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 4, 3))
    d = rng.standard_normal((2, 3, 2, 3))
    assert np.isclose((U.pool(x) * d).sum(), (x * U.pool_bwd(d)).sum())
    xs = rng.standard_normal((2, 3, 2, 3))
    ds = rng.standard_normal((2, 6, 4, 3))
    assert np.isclose((U.up(xs) * ds).sum(), (xs * U.up_bwd(ds)).sum())
    w = rng.standard_normal((3, 3, 3, 5))
    dout = rng.standard_normal((2, 6, 4, 5))
    dx, dw, db = U.conv3x3_bwd(x, w, dout)
    assert np.isclose((U.conv3x3(x, w, 0 * db) * dout).sum(), (x * dx).sum())
    # the convolution against a direct sum at one output pixel
    b, yy, xx, co = 1, 0, 3, 2
    s = sum(x[b, yy + ky - 1, xx + kx - 1, ci] * w[ky, kx, ci, co]
            for ky in range(3) for kx in range(3) for ci in range(3)
            if 0 <= yy + ky - 1 < 6 and 0 <= xx + kx - 1 < 4)
    assert np.isclose(U.conv3x3(x, w, np.zeros(5))[b, yy, xx, co], s)


def test_bf16_rounding():
    # bfloat16 spacing at 1 is 2^-7: 1 + 2^-8 and 1 + 3 2^-8 are ties (to even)
    a = np.array([1.0, 1.00390625, 1.01171875, 1.005859375, 3.0e-3, -2.5, 65504.0])
    r = U.to_bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625 and r[3] == 1.0078125
    assert r[5] == -2.5
    assert np.all(np.abs(r - a) <= np.abs(a) * 2 ** -8)
