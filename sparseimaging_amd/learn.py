"""Learned warm start for NUFFT-mode MFISTA.

A U-Net maps the dirty image Re(A^H W^2 V), a log-lambda plane and (third
channel) the dirty beam to a starting image; MFISTA then refines it to the
usual tolerance.  Cold-start MFISTA stays the fallback and decides which start
is kept: with verify=True a warm start whose converged objective is higher
than the cold start's is rejected.  The
reference has no learned component; this module serves its
mfista_imaging_core_nufft (c/mfista_nufft_lib.c:909) calls.

Network: L levels with c_l = base 2^l channels (default 4 levels, base 32).
Encoder levels 0 .. L-2: conv3x3 + ReLU twice, keep the skip, 2x2 average
pooling; bottleneck: conv3x3 + ReLU twice; decoder levels L-2 .. 0: nearest
2x upsampling, concatenation [upsampled, skip], conv3x3 + ReLU twice; head:
1x1 convolution to one channel, linear (a ReLU there dies on sparse targets);
the warm start clamps it at zero.
Weights are BF16 for the kernels with an FP32 master copy for Adam; every
convolution and its gradients run on the MFMA kernels of csrc/unet.hip
(include/mfista_learn.h).  torch only allocates device memory here.

Input normalisation: with s = max(dirty) / sum(w^2) (the peak flux the dirty
image implies), channel 0 is dirty / (s sum(w^2)) and channel 1 the constant
(log10 lambda_tsv - 2) / 2; the target is x / s and the warm start
alpha s max(y, 0), with alpha the exact minimiser of the objective along that
ray (one FP64 forward transform): MFISTA's iteration count is far more
sensitive to a wrong global scale than to local errors of the same size."""
import ctypes as C
import json
import os
import time

import numpy as np

from . import lib as _lib_loader
from . import ENGINE_PERSISTENT, MfistaError, PRESET_CXX12, mfista_imaging_core_nufft, mfista_imaging_core_nufft_batch
from . import nufft_adjoint, nufft_forward

_vp = C.c_void_p
_declared = False


def _lib():
    global _declared
    L = _lib_loader()
    if not _declared:
        i, l, f = C.c_int, C.c_long, C.c_float
        for n, args in {
            "unet_conv3x3_fwd": [_vp, i, _vp, i, _vp, _vp, _vp, i, i, i, i, i, _vp],
            "unet_conv3x3_bwd_data": [_vp, i, _vp, _vp, i, _vp, i, i, i, i, _vp],
            "unet_conv3x3_bwd_weight": [_vp, i, _vp, i, _vp, i, _vp, _vp, i, i, i, _vp],
            "unet_relu_bwd": [_vp, _vp, l, _vp],
            "unet_avgpool2_fwd": [_vp, _vp, i, i, i, i, _vp],
            "unet_avgpool2_bwd": [_vp, _vp, i, i, i, i, i, _vp],
            "unet_upsample2_fwd": [_vp, _vp, i, i, i, i, _vp],
            "unet_upsample2_bwd": [_vp, _vp, i, i, i, i, _vp],
            "unet_head_fwd": [_vp, i, _vp, _vp, _vp, l, i, _vp],
            "unet_head_bwd": [_vp, i, _vp, _vp, _vp, _vp, _vp, l, _vp],
            "unet_mse": [_vp, _vp, _vp, _vp, l, _vp],
            "unet_mse_weighted": [_vp, _vp, _vp, l, _vp, _vp, l, _vp],
            "unet_adam": [_vp, _vp, _vp, _vp, _vp, l, f, f, f, f, i, _vp],
            "unet_to_bf16": [_vp, _vp, l, _vp],
        }.items():
            fn = getattr(L, n)
            fn.restype = i
            fn.argtypes = args
        _declared = True
    return L


def _ck(rc):
    if rc != 0:
        raise MfistaError(rc)


def param_shapes(in_ch=2, base=32, levels=4):
    """Ordered (name, shape) of every parameter; 3x3 weights [ky][kx][ci][co]."""
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


class UNet:
    """The warm-start network on one GPU (torch tensors as device memory)."""

    def __init__(self, base=32, levels=4, in_ch=2, seed=0, device="cuda"):
        import torch

        self.torch = torch
        self.base, self.levels, self.in_ch = base, levels, in_ch
        self.dev = torch.device(device)
        self.shapes = param_shapes(in_ch, base, levels)
        rng = np.random.default_rng(seed)
        init = {}
        for name, shape in self.shapes:
            if name.endswith(".b"):
                init[name] = np.full(shape, 0.01)
            else:
                fan = int(np.prod(shape[:-1])) if len(shape) > 1 else shape[0]
                init[name] = rng.standard_normal(shape) * np.sqrt(2.0 / fan)
        self._alloc()
        self.set_params(init)
        self.step_count = 0

    def _alloc(self):
        t = self.torch
        self.w32, self.w16, self.g, self.m, self.v = {}, {}, {}, {}, {}
        for name, shape in self.shapes:
            n = int(np.prod(shape))
            self.w32[name] = t.zeros(n, dtype=t.float32, device=self.dev)
            self.w16[name] = t.zeros(n, dtype=t.int16, device=self.dev)
            self.g[name] = t.zeros(n, dtype=t.float32, device=self.dev)
            self.m[name] = t.zeros(n, dtype=t.float32, device=self.dev)
            self.v[name] = t.zeros(n, dtype=t.float32, device=self.dev)
        self.loss_buf = t.zeros(1, dtype=t.float64, device=self.dev)

    def _stream(self):
        return _vp(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def set_params(self, params):
        """Loads FP32 master weights from a dict of arrays (any float dtype)."""
        L = _lib()
        for name, shape in self.shapes:
            a = np.asarray(params[name], dtype=np.float32).reshape(-1)
            self.w32[name].copy_(self.torch.from_numpy(a))
            _ck(L.unet_to_bf16(self.w32[name].data_ptr(), self.w16[name].data_ptr(), a.size, self._stream()))

    def get_params(self):
        return {name: self.w32[name].cpu().numpy().reshape(shape) for name, shape in self.shapes}

    def save(self, path, meta=None):
        p = self.get_params()
        p["__meta__"] = np.frombuffer(json.dumps(dict(meta or {}, base=self.base, levels=self.levels,
                                                      in_ch=self.in_ch)).encode(), dtype=np.uint8)
        np.savez_compressed(path, **p)

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(d["__meta__"]).decode())
        net = cls(base=meta["base"], levels=meta["levels"], in_ch=meta["in_ch"], device=device)
        net.set_params({k: d[k] for k in d.files if k != "__meta__"})
        net.meta = meta
        return net

    # ------------------------------------------------------------------ passes
    def _empty(self, *shape):
        return self.torch.empty(*shape, dtype=self.torch.float32, device=self.dev)

    def forward(self, x):
        """x: float32 device tensor [B, H, W, in_ch] (NHWC); returns y [B, H, W]."""
        L, st = _lib(), self._stream()
        B, H, W, _ = x.shape
        c = [self.base * 2 ** l for l in range(self.levels)]
        cache = {"x": x}
        h, ch, hh, ww = x, self.in_ch, H, W
        for l in range(self.levels - 1):
            a = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_fwd(h.data_ptr(), ch, None, 0, self.w16[f"e{l}a.w"].data_ptr(),
                                   self.w32[f"e{l}a.b"].data_ptr(), a.data_ptr(), B, hh, ww, c[l], 1, st))
            b = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_fwd(a.data_ptr(), c[l], None, 0, self.w16[f"e{l}b.w"].data_ptr(),
                                   self.w32[f"e{l}b.b"].data_ptr(), b.data_ptr(), B, hh, ww, c[l], 1, st))
            cache[f"e{l}"] = (h, ch, a, b, hh, ww)
            p = self._empty(B, hh // 2, ww // 2, c[l])
            _ck(L.unet_avgpool2_fwd(b.data_ptr(), p.data_ptr(), B, hh, ww, c[l], st))
            h, ch, hh, ww = p, c[l], hh // 2, ww // 2
        cm = c[-1]
        a = self._empty(B, hh, ww, cm)
        _ck(L.unet_conv3x3_fwd(h.data_ptr(), ch, None, 0, self.w16["m.a.w"].data_ptr(), self.w32["m.a.b"].data_ptr(),
                               a.data_ptr(), B, hh, ww, cm, 1, st))
        b = self._empty(B, hh, ww, cm)
        _ck(L.unet_conv3x3_fwd(a.data_ptr(), cm, None, 0, self.w16["m.b.w"].data_ptr(), self.w32["m.b.b"].data_ptr(),
                               b.data_ptr(), B, hh, ww, cm, 1, st))
        cache["m"] = (h, ch, a, b, hh, ww)
        h, ch = b, cm
        for l in range(self.levels - 2, -1, -1):
            u = self._empty(B, 2 * hh, 2 * ww, ch)
            _ck(L.unet_upsample2_fwd(h.data_ptr(), u.data_ptr(), B, hh, ww, ch, st))
            hh, ww = 2 * hh, 2 * ww
            skip = cache[f"e{l}"][3]
            a = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_fwd(u.data_ptr(), ch, skip.data_ptr(), c[l], self.w16[f"d{l}a.w"].data_ptr(),
                                   self.w32[f"d{l}a.b"].data_ptr(), a.data_ptr(), B, hh, ww, c[l], 1, st))
            b = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_fwd(a.data_ptr(), c[l], None, 0, self.w16[f"d{l}b.w"].data_ptr(),
                                   self.w32[f"d{l}b.b"].data_ptr(), b.data_ptr(), B, hh, ww, c[l], 1, st))
            cache[f"d{l}"] = (u, ch, a, b, hh, ww)
            h, ch = b, c[l]
        y = self._empty(B, H, W)
        _ck(L.unet_head_fwd(h.data_ptr(), ch, self.w16["head.w"].data_ptr(), self.w32["head.b"].data_ptr(),
                            y.data_ptr(), B * H * W, 0, st))
        cache["head"] = (h, ch)
        cache["y"] = y
        self.cache = cache
        return y

    def zero_grad(self):
        for name in self.g:
            self.g[name].zero_()

    def backward(self, dy):
        """Adds the parameter gradients for dL/dy = dy ([B, H, W], consumed)."""
        L, st = _lib(), self._stream()
        cache = self.cache
        y = cache["y"]
        B, H, W = y.shape
        c = [self.base * 2 ** l for l in range(self.levels)]
        h, ch = cache["head"]
        dh = self._empty(B, H, W, ch)
        _ck(L.unet_head_bwd(h.data_ptr(), ch, self.w16["head.w"].data_ptr(), dy.data_ptr(), dh.data_ptr(),
                            self.g["head.w"].data_ptr(), self.g["head.b"].data_ptr(), B * H * W, st))
        dskip = {}
        for l in range(0, self.levels - 1):
            u, cu, a, b, hh, ww = cache[f"d{l}"]
            _ck(L.unet_relu_bwd(b.data_ptr(), dh.data_ptr(), dh.numel(), st))
            _ck(L.unet_conv3x3_bwd_weight(a.data_ptr(), c[l], None, 0, dh.data_ptr(), c[l],
                                          self.g[f"d{l}b.w"].data_ptr(), self.g[f"d{l}b.b"].data_ptr(), B, hh, ww,
                                          st))
            da = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_bwd_data(dh.data_ptr(), c[l], self.w16[f"d{l}b.w"].data_ptr(), da.data_ptr(), c[l],
                                        None, 0, B, hh, ww, st))
            _ck(L.unet_relu_bwd(a.data_ptr(), da.data_ptr(), da.numel(), st))
            skip = cache[f"e{l}"][3]
            _ck(L.unet_conv3x3_bwd_weight(u.data_ptr(), cu, skip.data_ptr(), c[l], da.data_ptr(), c[l],
                                          self.g[f"d{l}a.w"].data_ptr(), self.g[f"d{l}a.b"].data_ptr(), B, hh, ww,
                                          st))
            du = self._empty(B, hh, ww, cu)
            ds = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_bwd_data(da.data_ptr(), c[l], self.w16[f"d{l}a.w"].data_ptr(), du.data_ptr(), cu,
                                        ds.data_ptr(), c[l], B, hh, ww, st))
            dskip[l] = ds
            dh = self._empty(B, hh // 2, ww // 2, cu)
            _ck(L.unet_upsample2_bwd(du.data_ptr(), dh.data_ptr(), B, hh // 2, ww // 2, cu, st))
        hin, cin, a, b, hh, ww = cache["m"]
        cm = c[-1]
        _ck(L.unet_relu_bwd(b.data_ptr(), dh.data_ptr(), dh.numel(), st))
        _ck(L.unet_conv3x3_bwd_weight(a.data_ptr(), cm, None, 0, dh.data_ptr(), cm, self.g["m.b.w"].data_ptr(),
                                      self.g["m.b.b"].data_ptr(), B, hh, ww, st))
        da = self._empty(B, hh, ww, cm)
        _ck(L.unet_conv3x3_bwd_data(dh.data_ptr(), cm, self.w16["m.b.w"].data_ptr(), da.data_ptr(), cm, None, 0, B,
                                    hh, ww, st))
        _ck(L.unet_relu_bwd(a.data_ptr(), da.data_ptr(), da.numel(), st))
        _ck(L.unet_conv3x3_bwd_weight(hin.data_ptr(), cin, None, 0, da.data_ptr(), cm, self.g["m.a.w"].data_ptr(),
                                      self.g["m.a.b"].data_ptr(), B, hh, ww, st))
        dh = self._empty(B, hh, ww, cin)
        _ck(L.unet_conv3x3_bwd_data(da.data_ptr(), cm, self.w16["m.a.w"].data_ptr(), dh.data_ptr(), cin, None, 0, B,
                                    hh, ww, st))
        for l in range(self.levels - 2, -1, -1):
            hin, cin, a, b, hh, ww = cache[f"e{l}"]
            db = dskip[l]
            _ck(L.unet_avgpool2_bwd(dh.data_ptr(), db.data_ptr(), B, hh, ww, c[l], 1, st))
            _ck(L.unet_relu_bwd(b.data_ptr(), db.data_ptr(), db.numel(), st))
            _ck(L.unet_conv3x3_bwd_weight(a.data_ptr(), c[l], None, 0, db.data_ptr(), c[l],
                                          self.g[f"e{l}b.w"].data_ptr(), self.g[f"e{l}b.b"].data_ptr(), B, hh, ww,
                                          st))
            da = self._empty(B, hh, ww, c[l])
            _ck(L.unet_conv3x3_bwd_data(db.data_ptr(), c[l], self.w16[f"e{l}b.w"].data_ptr(), da.data_ptr(), c[l],
                                        None, 0, B, hh, ww, st))
            _ck(L.unet_relu_bwd(a.data_ptr(), da.data_ptr(), da.numel(), st))
            _ck(L.unet_conv3x3_bwd_weight(hin.data_ptr(), cin, None, 0, da.data_ptr(), c[l],
                                          self.g[f"e{l}a.w"].data_ptr(), self.g[f"e{l}a.b"].data_ptr(), B, hh, ww,
                                          st))
            if l > 0:
                dh = self._empty(B, hh, ww, cin)
                _ck(L.unet_conv3x3_bwd_data(da.data_ptr(), c[l], self.w16[f"e{l}a.w"].data_ptr(), dh.data_ptr(), cin,
                                            None, 0, B, hh, ww, st))

    def loss(self, y, t, want_grad=True):
        """Mean squared error (float) and, if want_grad, its gradient (a new tensor)."""
        L = _lib()
        dy = self._empty(*y.shape) if want_grad else None
        _ck(L.unet_mse(y.data_ptr(), t.data_ptr(), dy.data_ptr() if dy is not None else None,
                       self.loss_buf.data_ptr(), y.numel(), self._stream()))
        return float(self.loss_buf.item()), dy

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        L, st = _lib(), self._stream()
        self.step_count += 1
        for name in self.w32:
            _ck(L.unet_adam(self.w32[name].data_ptr(), self.w16[name].data_ptr(), self.g[name].data_ptr(),
                            self.m[name].data_ptr(), self.v[name].data_ptr(), self.w32[name].numel(), lr, beta1,
                            beta2, eps, self.step_count, st))

    def loss_weighted(self, y, t, wt):
        """Per-sample weighted squared error: sum_b wt[b] |y_b - t_b|^2 / numel and its gradient."""
        L = _lib()
        dy = self._empty(*y.shape)
        _ck(L.unet_mse_weighted(y.data_ptr(), t.data_ptr(), wt.data_ptr(), y[0].numel(), dy.data_ptr(),
                                self.loss_buf.data_ptr(), y.numel(), self._stream()))
        return float(self.loss_buf.item()), dy

    def train_step(self, x, t, lr, wt=None):
        """One Adam step on a batch; wt (per-sample weights, device float32)
        selects the weighted loss."""
        self.zero_grad()
        y = self.forward(x)
        loss, dy = self.loss(y, t) if wt is None else self.loss_weighted(y, t, wt)
        self.backward(dy)
        self.adam(lr)
        return loss


# ---------------------------------------------------------------------- data
def net_input(u, v, vis_r, vis_i, vis_std, N, lambda_tsv, preset=PRESET_CXX12, channels=2):
    """(N, N, channels) network input and the flux scale s (see the module
    docstring).  channels=3 adds the dirty beam Re(A^H W^2 1) / sum(w^2) (the
    point-spread function of this (u,v) coverage, peak 1 at the centre): the
    coverage differs from problem to problem, and without it the network
    cannot know what the dirty image is convolved with."""
    w2 = 1.0 / np.asarray(vis_std) ** 2
    dirty = nufft_adjoint(u, v, w2 * (np.asarray(vis_r) + 1j * np.asarray(vis_i)), N, N, preset=preset, fp64=True)
    sw = float(w2.sum())
    peak = float(dirty.max())
    s = max(peak, 1e-30) / sw
    x = np.empty((N, N, channels), dtype=np.float32)
    x[..., 0] = dirty / (s * sw)
    x[..., 1] = (np.log10(lambda_tsv) - 2.0) / 2.0
    if channels == 3:
        x[..., 2] = nufft_adjoint(u, v, w2.astype(np.complex128), N, N, preset=preset, fp64=True) / sw
    return x, s


def random_sky(N, rng):
    """A ring (random radius, width and centre) plus 1-4 Gaussians, peak 1."""
    from .synth import gaussian_blobs, ring

    c = N / 2 + rng.uniform(-0.08, 0.08, size=2) * N
    x = ring(N, rng.uniform(0.08, 0.2) * N, rng.uniform(0.015, 0.04) * N, c[0], c[1])
    x = x * rng.uniform(0.3, 1.0) + gaussian_blobs(N, rng, int(rng.integers(1, 5)))
    x = np.maximum(x, 0.0)
    return x / x.max()


def make_problem(N, M, rng, preset=PRESET_CXX12, noise_frac=0.01):
    """A random sky observed on 12 random elliptical arcs; visibilities from the
    FP64 GPU NUFFT of the same preset plus complex noise."""
    from .synth import elliptical_arcs

    x = random_sky(N, rng)
    u, v = elliptical_arcs(M, rng)
    vis = nufft_forward(u, v, x, preset=preset, fp64=True)
    sig = noise_frac * np.abs(vis).max()
    vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    return {"u": u, "v": v, "vis_r": vis.real.copy(), "vis_i": vis.imag.copy(), "std": np.full(M, sig),
            "x_true": x}


def make_training_set(n, N=256, M=20000, seed=5000, lambda_l1=1.0, log10_tsv=(0.0, 4.0), batch=32,
                      rel_tol=1e-6, preset=PRESET_CXX12, log=None, channels=2, engine=ENGINE_PERSISTENT):
    """n (input, target) pairs: random skies solved by the GPU MFISTA (cold
    start, `rel_tol` relative tolerance) with lambda_tsv log-uniform; `batch`
    problems per launch (1: single solves on the persistent engine; more: one
    batch per launch on `engine`, ENGINE_STREAM being the wide engine)."""
    rng = np.random.default_rng(seed)
    X = np.empty((n, N, N, channels), dtype=np.float32)
    T = np.empty((n, N, N), dtype=np.float32)
    lam = np.empty(n)
    t0 = time.time()
    for k0 in range(0, n, batch):
        probs, ss = [], []
        for k in range(k0, min(n, k0 + batch)):
            P = make_problem(N, M, rng, preset)
            lt = 10.0 ** rng.uniform(*log10_tsv)
            lam[k] = lt
            X[k], s = net_input(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], N, lt, preset, channels)
            ss.append(s)
            probs.append(dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"],
                              NX=N, NY=N, maxiter=50000, eps=0.0, lambda_l1=lambda_l1, lambda_tsv=lt, cinit=1e4,
                              nonneg_flag=1))
        if len(probs) == 1:  # one solve uses the whole GPU
            q = dict(probs[0])
            args = [q.pop(k) for k in ("u_dx", "v_dy", "vis_r", "vis_i", "vis_std", "NX", "NY")]
            xs = [mfista_imaging_core_nufft(*args, preset=preset, rel_tol=rel_tol, **q)[0]]
        else:
            xs, _, _ = mfista_imaging_core_nufft_batch(probs, preset=preset, rel_tol=rel_tol, engine=engine)
        for j, xk in enumerate(xs):
            T[k0 + j] = xk / ss[j]
        if log and ((k0 // batch) % max(1, 64 // batch) == 0 or k0 + batch >= n):
            log(f"training set: {min(n, k0 + batch)}/{n} solved ({time.time() - t0:.1f} s)")
    return X, T, lam


def train(net, X, T, epochs=10, batch=8, lr=1e-3, lr_final=1e-4, log=None, relative=False):
    """Adam with a geometric learning-rate decay; returns the loss history.
    relative=True weights sample b by numel / |T_b|^2, so the loss is the mean
    squared relative error: MFISTA's iteration count depends on the start's
    relative distance to the solution, and plain MSE on mostly-zero images
    rewards a blurred average."""
    torch = net.torch
    n = X.shape[0]
    steps = epochs * ((n + batch - 1) // batch)
    hist = []
    rng = np.random.default_rng(0)
    k = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = order[b0:b0 + batch]
            x = torch.from_numpy(X[idx]).to(net.dev)
            t = torch.from_numpy(T[idx]).to(net.dev)
            cur = lr * (lr_final / lr) ** (k / max(steps - 1, 1))
            wt = None
            if relative:
                tn = np.maximum((T[idx].reshape(len(idx), -1).astype(np.float64) ** 2).sum(1), 1e-30)
                wt = torch.from_numpy((T[idx][0].size / tn).astype(np.float32)).to(net.dev)
            tot += net.train_step(x, t, cur, wt) * len(idx)
            k += 1
        hist.append(tot / n)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}")
    return hist


def predict(net, x_in, s):
    """Warm start s max(y, 0) for one (N, N, 2) input."""
    torch = net.torch
    y = net.forward(torch.from_numpy(x_in[None]).to(net.dev))
    return np.maximum(y[0].cpu().numpy().astype(np.float64), 0.0) * s


def objective(u, v, vis_r, vis_i, vis_std, x, lambda_l1, lambda_tsv, preset=PRESET_CXX12):
    """The L1+TSV objective of an NX x NY image x (FP64 NUFFT), as RESULT.finalcost
    (c/mfista_nufft_lib.c:791-905): |W(V - A x)|^2 / 2 + lambda_l1 |x|_1 + lambda_tsv TSV(x)."""
    x = np.asarray(x, dtype=np.float64)
    m = nufft_forward(u, v, x, preset=preset, fp64=True)
    r = ((np.asarray(vis_r) + 1j * np.asarray(vis_i)) - m) / np.asarray(vis_std)
    tsv = float(np.sum(np.diff(x, axis=0) ** 2) + np.sum(np.diff(x, axis=1) ** 2))
    return 0.5 * float(np.sum(np.abs(r) ** 2)) + lambda_l1 * float(np.abs(x).sum()) + lambda_tsv * tsv


def ray_scale(u, v, vis_r, vis_i, vis_std, x0, lambda_l1, lambda_tsv, preset=PRESET_CXX12):
    """alpha >= 0 minimising F(alpha x0), a quadratic in alpha:
    F = |W(V - alpha A x0)|^2 / 2 + alpha lambda_l1 |x0|_1 + alpha^2 lambda_tsv TSV(x0), so
    alpha = (Re<W A x0, W V> - lambda_l1 |x0|_1) / (|W A x0|^2 + 2 lambda_tsv TSV(x0)).
    One FP64 forward transform; corrects the global flux scale of a prediction,
    to which MFISTA's iteration count is most sensitive."""
    x0 = np.asarray(x0, dtype=np.float64)
    w = 1.0 / np.asarray(vis_std)
    m = nufft_forward(u, v, x0, preset=preset, fp64=True) * w
    V = (np.asarray(vis_r) + 1j * np.asarray(vis_i)) * w
    tsv = float(np.sum(np.diff(x0, axis=0) ** 2) + np.sum(np.diff(x0, axis=1) ** 2))
    den = float(np.sum(np.abs(m) ** 2)) + 2.0 * lambda_tsv * tsv
    if den <= 0.0:
        return 0.0
    return max(0.0, (float(np.real(np.vdot(m, V))) - lambda_l1 * float(np.abs(x0).sum())) / den)


def solve_warm(net, u, v, vis_r, vis_i, vis_std, N, lambda_l1=1.0, lambda_tsv=100.0, verify=False,
               preset=PRESET_CXX12, **kw):
    """MFISTA from the network's warm start, rescaled by the exact line search
    along its ray (ray_scale).  With verify=True the cold start is solved too
    and wins if its converged objective is lower (the acceptance rule).
    Returns (x, RESULT dict, info) with info["warm_accepted"], info["alpha"]
    and info["x0"] (the starting image)."""
    x_in, s = net_input(u, v, vis_r, vis_i, vis_std, N, lambda_tsv, preset, net.in_ch)
    t0 = time.perf_counter()
    x0 = predict(net, x_in, s)
    alpha = ray_scale(u, v, vis_r, vis_i, vis_std, x0, lambda_l1, lambda_tsv, preset)
    x0 = alpha * x0
    t_net = time.perf_counter() - t0
    xw, rw, iw = mfista_imaging_core_nufft(u, v, vis_r, vis_i, vis_std, N, N, xinit=x0, lambda_l1=lambda_l1,
                                           lambda_tsv=lambda_tsv, preset=preset, **kw)
    info = dict(iw, net_ms=t_net * 1e3, warm_accepted=True, warm_iter=rw["ITER"], alpha=alpha, x0=x0)
    if verify:
        xc, rc, ic = mfista_imaging_core_nufft(u, v, vis_r, vis_i, vis_std, N, N, lambda_l1=lambda_l1,
                                               lambda_tsv=lambda_tsv, preset=preset, **kw)
        info.update(cold_iter=rc["ITER"], cold_finalcost=rc["finalcost"])
        if rw["finalcost"] > rc["finalcost"]:
            info["warm_accepted"] = False
            return xc, rc, info
    return xw, rw, info


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_256.npz")
