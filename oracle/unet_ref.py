"""FP64 numpy restatement of the learned warm start's network -- test
infrastructure, not product code: it is the checker of the HIP kernels in
sparseimaging_amd/csrc/unet.hip (tests only).

There is no reference counterpart (the reference has no learned component);
this module is pinned by finite differences (tests/test_oracle_unet.py).

Network (the same description as sparseimaging_amd/learn.py): L levels with
c_l = base * 2**l channels.  Encoder levels 0 .. L-2: conv3x3 + ReLU twice,
keep the skip, 2x2 average pooling; bottleneck: conv3x3 + ReLU twice at level
L-1; decoder levels L-2 .. 0: nearest 2x upsampling, concatenation
[upsampled, skip], conv3x3 + ReLU twice; head: 1x1 convolution to one
channel (linear; the warm start clamps it at zero).  Tensors are NHWC, 3x3
weights [ky][kx][ci][co].

bf16=True in forward/backward rounds every convolution operand (the
activations entering a convolution, the gradients entering its backward) to
bfloat16, as the GPU's matrix-core kernels do; with bf16-valued weights the
result then differs from the GPU only by FP32 accumulation order."""
import numpy as np


def param_shapes(in_ch=2, base=32, levels=4):
    """Ordered (name, shape) of every parameter."""
    c = [base * 2 ** l for l in range(levels)]
    shapes = []
    prev = in_ch
    for l in range(levels - 1):
        shapes += [(f"e{l}a.w", (3, 3, prev, c[l])), (f"e{l}a.b", (c[l],)),
                   (f"e{l}b.w", (3, 3, c[l], c[l])), (f"e{l}b.b", (c[l],))]
        prev = c[l]
    L = levels - 1
    shapes += [("m.a.w", (3, 3, prev, c[L])), ("m.a.b", (c[L],)), ("m.b.w", (3, 3, c[L], c[L])), ("m.b.b", (c[L],))]
    prev = c[L]
    for l in range(levels - 2, -1, -1):
        shapes += [(f"d{l}a.w", (3, 3, prev + c[l], c[l])), (f"d{l}a.b", (c[l],)),
                   (f"d{l}b.w", (3, 3, c[l], c[l])), (f"d{l}b.b", (c[l],))]
        prev = c[l]
    shapes += [("head.w", (prev,)), ("head.b", (1,))]
    return shapes


def conv3x3(x, w, b):
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, H, W, w.shape[3]))
    for ky in range(3):
        for kx in range(3):
            out += xp[:, ky:ky + H, kx:kx + W, :] @ w[ky, kx]
    return out + b


def conv3x3_bwd(x, w, dout, dout_bias=None):
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + H, kx:kx + W, :] += dout @ w[ky, kx].T
            dw[ky, kx] = xp[:, ky:ky + H, kx:kx + W, :].reshape(-1, x.shape[3]).T @ dout.reshape(-1, w.shape[3])
    db = (dout if dout_bias is None else dout_bias).sum(axis=(0, 1, 2))
    return dxp[:, 1:-1, 1:-1, :], dw, db


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(params, x, levels=4, bf16=False):
    """Returns (y[B,H,W,1], cache)."""
    r = to_bf16 if bf16 else (lambda a: a)
    cache = {"x": x, "bf16": bf16}
    h = x
    skips = []
    for l in range(levels - 1):
        a = np.maximum(conv3x3(r(h), params[f"e{l}a.w"], params[f"e{l}a.b"]), 0)
        bb = np.maximum(conv3x3(r(a), params[f"e{l}b.w"], params[f"e{l}b.b"]), 0)
        cache[f"e{l}in"], cache[f"e{l}a"], cache[f"e{l}b"] = h, a, bb
        skips.append(bb)
        h = pool(bb)
    a = np.maximum(conv3x3(r(h), params["m.a.w"], params["m.a.b"]), 0)
    bb = np.maximum(conv3x3(r(a), params["m.b.w"], params["m.b.b"]), 0)
    cache["m.in"], cache["m.a"], cache["m.b"] = h, a, bb
    h = bb
    for l in range(levels - 2, -1, -1):
        u = up(h)
        cat = np.concatenate([u, skips[l]], axis=3)
        a = np.maximum(conv3x3(r(cat), params[f"d{l}a.w"], params[f"d{l}a.b"]), 0)
        bb = np.maximum(conv3x3(r(a), params[f"d{l}b.w"], params[f"d{l}b.b"]), 0)
        cache[f"d{l}cat"], cache[f"d{l}a"], cache[f"d{l}b"] = cat, a, bb
        h = bb
    y = (h @ params["head.w"] + params["head.b"][0])[..., None]
    cache["head.in"], cache["y"] = h, y
    return y, cache


def backward(params, cache, dy, levels=4):
    """Gradients of every parameter for dL/dy = dy (operands rounded to
    bfloat16 like the forward pass if it was run with bf16=True)."""
    r = to_bf16 if cache.get("bf16") else (lambda a: a)
    g = {}
    dz = dy[..., 0]
    h = cache["head.in"]
    g["head.w"] = (h * dz[..., None]).sum(axis=(0, 1, 2))
    g["head.b"] = np.array([dz.sum()])
    dh = dz[..., None] * params["head.w"]
    dskips = {}
    # decoder, from level 0 up to level L-2 (the reverse of the forward order)
    for l in range(0, levels - 1):
        bb = cache[f"d{l}b"]
        dh = dh * (bb > 0)
        da, g[f"d{l}b.w"], g[f"d{l}b.b"] = conv3x3_bwd(r(cache[f"d{l}a"]), params[f"d{l}b.w"], r(dh), dh)
        da = da * (cache[f"d{l}a"] > 0)
        dcat, g[f"d{l}a.w"], g[f"d{l}a.b"] = conv3x3_bwd(r(cache[f"d{l}cat"]), params[f"d{l}a.w"], r(da), da)
        cu = dcat.shape[3] - cache[f"e{l}b"].shape[3]
        dskips[l] = dcat[..., cu:]
        dh = up_bwd(dcat[..., :cu])
    bb = cache["m.b"]
    dh = dh * (bb > 0)
    da, g["m.b.w"], g["m.b.b"] = conv3x3_bwd(r(cache["m.a"]), params["m.b.w"], r(dh), dh)
    da = da * (cache["m.a"] > 0)
    dh, g["m.a.w"], g["m.a.b"] = conv3x3_bwd(r(cache["m.in"]), params["m.a.w"], r(da), da)
    for l in range(levels - 2, -1, -1):
        dbb = pool_bwd(dh) + dskips[l]
        dbb = dbb * (cache[f"e{l}b"] > 0)
        da, g[f"e{l}b.w"], g[f"e{l}b.b"] = conv3x3_bwd(r(cache[f"e{l}a"]), params[f"e{l}b.w"], r(dbb), dbb)
        da = da * (cache[f"e{l}a"] > 0)
        dh, g[f"e{l}a.w"], g[f"e{l}a.b"] = conv3x3_bwd(r(cache[f"e{l}in"]), params[f"e{l}a.w"], r(da), da)
    return g


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def init_params(rng, in_ch=2, base=32, levels=4, bf16=True):
    """He-initialised parameters; with bf16=True every weight is rounded to
    bfloat16 (so the GPU's BF16 weight copy holds exactly the same values)."""
    p = {}
    for name, shape in param_shapes(in_ch, base, levels):
        if name.endswith(".b"):
            p[name] = np.full(shape, 0.01)
        else:
            fan = int(np.prod(shape[:-1])) if len(shape) > 1 else shape[0]
            w = rng.standard_normal(shape) * np.sqrt(2.0 / fan)
            p[name] = to_bf16(w) if bf16 else w
    return p


def to_bf16(a):
    """Round to the nearest bfloat16 (ties to even) and back to float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)
