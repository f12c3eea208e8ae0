"""The learned warm start's GPU kernels and network against the FP64 numpy
network (oracle/unet_ref.py, pinned by finite differences), its training,
and the warm-start acceptance rule.

Kernel bars: the MFMA kernels round their operands to bfloat16 and
accumulate in FP32; the single-layer checks give the reference the same
bfloat16 operands, so only the accumulation order differs (relative L2 <=
1e-5).  The whole-network check runs the numpy network with the same
bfloat16 operand rounding (relative L2 <= 1e-3 for the output and every
gradient; the rounding points are the same, the FP32 accumulation order is
not) and reports the distance to the unrounded FP64 network."""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U
import sparseimaging_amd as sa
from sparseimaging_amd import learn

pytestmark = pytest.mark.gpu
dev = torch.device("cuda:0")


def _t(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(dev)


def _bf16_bits(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return torch.from_numpy(u.view(np.int16).copy()).to(dev)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", [(2, 12, 10, 5, 3, 70), (2, 12, 10, 32, 32, 40), (3, 9, 11, 64, 32, 64)],
                         ids=["gather", "vec-fwd-wgrad", "vec-all"])
def test_conv3x3_layers_match_oracle(shape):
    """Channel counts that are multiples of 32 take the vectorised staging
    path (8 channels per load); the others the element gather.  Pixel counts
    are not multiples of the 64-row tile in any case."""
    rng = np.random.default_rng(1)
    B, H, W, C1, C2, Co = shape
    x1 = rng.standard_normal((B, H, W, C1))
    x2 = rng.standard_normal((B, H, W, C2))
    w = U.to_bf16(rng.standard_normal((3, 3, C1 + C2, Co)) * 0.3)
    b = rng.standard_normal(Co) * 0.1
    L = learn._lib()
    st = learn._vp(torch.cuda.current_stream(dev).cuda_stream)
    out = torch.empty((B, H, W, Co), dtype=torch.float32, device=dev)
    g1, g2, gw16 = _t(x1), _t(x2), _bf16_bits(w)
    gb = _t(b)
    print(f"case conv3x3 forward / input grad / weight grad, B={B} {H}x{W}, Cin={C1}+{C2}, Cout={Co}", flush=True)
    assert L.unet_conv3x3_fwd(g1.data_ptr(), C1, g2.data_ptr(), C2, gw16.data_ptr(), gb.data_ptr(), out.data_ptr(),
                              B, H, W, Co, 0, st) == 0
    xc = np.concatenate([U.to_bf16(x1), U.to_bf16(x2)], axis=3)
    ref = U.conv3x3(xc, w, b)
    e = _rel(out.cpu().numpy(), ref)
    print(f"  forward relL2 {e:.2e}", flush=True)
    assert e <= 1e-5
    # relu variant
    assert L.unet_conv3x3_fwd(g1.data_ptr(), C1, g2.data_ptr(), C2, gw16.data_ptr(), gb.data_ptr(), out.data_ptr(),
                              B, H, W, Co, 1, st) == 0
    assert _rel(out.cpu().numpy(), np.maximum(ref, 0)) <= 1e-5
    dout = rng.standard_normal((B, H, W, Co))
    d1 = torch.empty((B, H, W, C1), dtype=torch.float32, device=dev)
    d2 = torch.empty((B, H, W, C2), dtype=torch.float32, device=dev)
    gd = _t(dout)
    assert L.unet_conv3x3_bwd_data(gd.data_ptr(), Co, gw16.data_ptr(), d1.data_ptr(), C1, d2.data_ptr(), C2, B, H, W,
                                   st) == 0
    dx, dw, db = U.conv3x3_bwd(xc, w, U.to_bf16(dout))
    e1, e2 = _rel(d1.cpu().numpy(), dx[..., :C1]), _rel(d2.cpu().numpy(), dx[..., C1:])
    print(f"  input grad relL2 {e1:.2e} {e2:.2e}", flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5
    gwt = torch.zeros(9 * (C1 + C2) * Co, dtype=torch.float32, device=dev)
    gbt = torch.zeros(Co, dtype=torch.float32, device=dev)
    assert L.unet_conv3x3_bwd_weight(g1.data_ptr(), C1, g2.data_ptr(), C2, gd.data_ptr(), Co, gwt.data_ptr(),
                                     gbt.data_ptr(), B, H, W, st) == 0
    _, dw_ref, _ = U.conv3x3_bwd(xc, w, U.to_bf16(dout))
    ew = _rel(gwt.cpu().numpy().reshape(dw_ref.shape), dw_ref)
    eb = _rel(gbt.cpu().numpy(), dout.sum(axis=(0, 1, 2)))
    print(f"  weight grad relL2 {ew:.2e}, bias grad {eb:.2e}", flush=True)
    assert ew <= 1e-5 and eb <= 1e-5


def test_pool_upsample_head_mse_adam():
    rng = np.random.default_rng(2)
    L = learn._lib()
    st = learn._vp(torch.cuda.current_stream(dev).cuda_stream)
    B, H, W, C = 2, 8, 6, 5
    x = rng.standard_normal((B, H, W, C))
    gx = _t(x)
    p = torch.empty((B, H // 2, W // 2, C), dtype=torch.float32, device=dev)
    assert L.unet_avgpool2_fwd(gx.data_ptr(), p.data_ptr(), B, H, W, C, st) == 0
    assert _rel(p.cpu().numpy(), U.pool(x)) <= 1e-6
    d = rng.standard_normal((B, H // 2, W // 2, C))
    din = _t(x)  # accumulate onto x
    assert L.unet_avgpool2_bwd(_t(d).data_ptr(), din.data_ptr(), B, H, W, C, 1, st) == 0
    assert _rel(din.cpu().numpy(), x + U.pool_bwd(d)) <= 1e-6
    u = torch.empty((B, 2 * H, 2 * W, C), dtype=torch.float32, device=dev)
    assert L.unet_upsample2_fwd(gx.data_ptr(), u.data_ptr(), B, H, W, C, st) == 0
    assert _rel(u.cpu().numpy(), U.up(x)) <= 1e-6
    du = rng.standard_normal((B, 2 * H, 2 * W, C))
    dd = torch.empty((B, H, W, C), dtype=torch.float32, device=dev)
    assert L.unet_upsample2_bwd(_t(du).data_ptr(), dd.data_ptr(), B, H, W, C, st) == 0
    assert _rel(dd.cpu().numpy(), U.up_bwd(du)) <= 1e-6
    # head: 1x1 to one channel
    hw = U.to_bf16(rng.standard_normal(C))
    hb = np.array([0.2])
    y = torch.empty((B, H, W), dtype=torch.float32, device=dev)
    assert L.unet_head_fwd(gx.data_ptr(), C, _bf16_bits(hw).data_ptr(), _t(hb).data_ptr(), y.data_ptr(), B * H * W, 0,
                           st) == 0
    assert _rel(y.cpu().numpy(), x @ hw + 0.2) <= 1e-6
    dy = rng.standard_normal((B, H, W))
    dxh = torch.empty((B, H, W, C), dtype=torch.float32, device=dev)
    dwh = torch.zeros(C, dtype=torch.float32, device=dev)
    dbh = torch.zeros(1, dtype=torch.float32, device=dev)
    assert L.unet_head_bwd(gx.data_ptr(), C, _bf16_bits(hw).data_ptr(), _t(dy).data_ptr(), dxh.data_ptr(),
                           dwh.data_ptr(), dbh.data_ptr(), B * H * W, st) == 0
    assert _rel(dxh.cpu().numpy(), dy[..., None] * hw) <= 1e-6
    assert _rel(dwh.cpu().numpy(), (x * dy[..., None]).sum(axis=(0, 1, 2))) <= 1e-5
    assert abs(dbh.item() - dy.sum()) <= 1e-4 * abs(dy).sum()
    # mse
    t = rng.standard_normal((B, H, W))
    loss = torch.zeros(1, dtype=torch.float64, device=dev)
    g = torch.empty((B, H, W), dtype=torch.float32, device=dev)
    assert L.unet_mse(_t(y.cpu().numpy()).data_ptr(), _t(t).data_ptr(), g.data_ptr(), loss.data_ptr(), B * H * W,
                      st) == 0
    lref, gref = U.mse(y.cpu().numpy().astype(np.float64), t)
    assert abs(loss.item() - lref) <= 1e-6 * lref and _rel(g.cpu().numpy(), gref) <= 1e-6
    # Adam step
    n = 1000
    w0, gg = rng.standard_normal(n), rng.standard_normal(n)
    w32, w16 = _t(w0), torch.zeros(n, dtype=torch.int16, device=dev)
    m, v = torch.zeros(n, dtype=torch.float32, device=dev), torch.zeros(n, dtype=torch.float32, device=dev)
    assert L.unet_adam(w32.data_ptr(), w16.data_ptr(), _t(gg).data_ptr(), m.data_ptr(), v.data_ptr(), n, 1e-2, 0.9,
                       0.999, 1e-8, 1, st) == 0
    # first step: m = 0.1 g, v = 0.001 g^2, update = lr sign(g) (up to eps)
    assert np.allclose(w32.cpu().numpy(), w0 - 1e-2 * np.sign(gg), rtol=0, atol=1e-6)
    bits = w16.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    assert np.allclose(bits.view(np.float32), U.to_bf16(w32.cpu().numpy()), rtol=0, atol=0)


def test_weighted_mse():
    """unet_mse_weighted: sum_b w_b |y_b - t_b|^2 / n and its gradient, against numpy."""
    rng = np.random.default_rng(4)
    B, per = 3, 50
    y, t = rng.standard_normal((B, per)), rng.standard_normal((B, per))
    w = np.array([0.5, 2.0, 1.5])
    L = learn._lib()
    st = learn._vp(torch.cuda.current_stream(dev).cuda_stream)
    gy, gt, gw = _t(y), _t(t), _t(w)
    dy = torch.empty_like(gy)
    loss = torch.zeros(1, dtype=torch.float64, device=dev)
    assert L.unet_mse_weighted(gy.data_ptr(), gt.data_ptr(), gw.data_ptr(), per, dy.data_ptr(), loss.data_ptr(),
                               B * per, st) == 0
    ref = float(np.sum(w[:, None] * (y - t) ** 2) / (B * per))
    print(f"case weighted MSE: {loss.item():.8e} vs {ref:.8e}", flush=True)
    assert abs(loss.item() - ref) <= 1e-6 * ref
    assert np.allclose(dy.cpu().numpy(), 2 * w[:, None] * (y - t) / (B * per), rtol=1e-5, atol=1e-8)


def test_network_forward_and_gradients_match_oracle():
    rng = np.random.default_rng(3)
    base, levels = 4, 3
    params = U.init_params(rng, 2, base, levels, bf16=True)
    net = learn.UNet(base=base, levels=levels, device="cuda:0")
    net.set_params(params)
    x = rng.standard_normal((2, 16, 16, 2))
    t = rng.standard_normal((2, 16, 16, 1)) * 0.5
    print(f"case U-Net forward + backward, base {base}, {levels} levels, B=2 16x16", flush=True)
    y = net.forward(_t(x))
    yr, cache = U.forward(params, x, levels, bf16=True)
    y64, _ = U.forward(params, x, levels)
    ey = _rel(y.cpu().numpy(), yr[..., 0])
    print(f"  output relL2 {ey:.2e} (vs the unrounded FP64 network {_rel(y.cpu().numpy(), y64[..., 0]):.2e})",
          flush=True)
    assert ey <= 1e-3
    net.zero_grad()
    loss, dy = net.loss(y, _t(t[..., 0]))
    net.backward(dy)
    lref, dyr = U.mse(yr, t)
    gref = U.backward(params, cache, dyr, levels)
    worst = 0.0
    dot_gpu = dot_ref = 0.0
    vrng = np.random.default_rng(4)
    for name, shape in learn.param_shapes(2, base, levels):
        gg = net.g[name].cpu().numpy().reshape(shape)
        e = _rel(gg, gref[name]) if np.linalg.norm(gref[name]) > 0 else float(np.linalg.norm(gg))
        worst = max(worst, e)
        vv = vrng.standard_normal(shape)
        dot_gpu += float((gg * vv).sum())
        dot_ref += float((gref[name] * vv).sum())
    dd = abs(dot_gpu - dot_ref) / abs(dot_ref)
    print(f"  loss {loss:.5e} (oracle {lref:.5e}); worst gradient relL2 {worst:.2e}; directional {dd:.2e}", flush=True)
    assert abs(loss - lref) <= 1e-3 * lref
    assert worst <= 1e-3
    assert dd <= 1e-3


def test_training_reduces_loss():
    rng = np.random.default_rng(5)
    net = learn.UNet(base=8, levels=3, seed=1, device="cuda:0")
    X = rng.standard_normal((8, 32, 32, 2)).astype(np.float32)
    T = np.maximum(X[..., 0] * 0.5 + 0.2, 0).astype(np.float32)  # a learnable map
    hist = learn.train(net, X, T, epochs=60, batch=4, lr=3e-3, lr_final=1e-3)
    print(f"case training 8 samples 32x32: loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.2 * hist[0]


def test_warm_start_gain_and_acceptance_rule():
    """Train a small network on GPU-solved 64^2 problems.  On held-out problems:
    the warm start (network output rescaled by the exact line search along its
    ray) must start closer to the optimum than the cold start -- on average
    less than half of the cold start's objective gap F(0) - F* -- and the
    acceptance rule must return the better of the two converged solves.  The
    iteration counts are reported; with a network this small they are not a
    reliable signal (the 256^2 network's gain is measured by bench.py)."""
    N, M = 64, 1500
    kw = dict(maxiter=50000, eps=0.0, cinit=1e4, nonneg_flag=1, rel_tol=1e-6)
    print("case warm start: 192 training problems 64x64, M=1500", flush=True)
    X, T, lam = learn.make_training_set(192, N=N, M=M, seed=900, batch=64, log=print)
    net = learn.UNet(base=16, levels=3, seed=2, device="cuda:0")
    hist = learn.train(net, X, T, epochs=25, batch=8, lr=2e-3, lr_final=2e-4)
    print(f"  training loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    rng = np.random.default_rng(77)
    wi, ci, gaps = [], [], []
    for k in range(8):
        P = learn.make_problem(N, M, rng)
        lt = 10.0 ** rng.uniform(0, 4)
        x, r, info = learn.solve_warm(net, P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], N, lambda_l1=1.0,
                                      lambda_tsv=lt, verify=True, **kw)
        best = min(info["cold_finalcost"], r["finalcost"])
        assert r["finalcost"] == best
        assert (info["warm_accepted"] and r["ITER"] == info["warm_iter"]) or not info["warm_accepted"]
        fs = best
        f0 = learn.objective(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], np.zeros((N, N)), 1.0, lt)
        fw = learn.objective(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], info["x0"], 1.0, lt)
        gaps.append((fw - fs) / (f0 - fs))
        wi.append(info["warm_iter"])
        ci.append(info["cold_iter"])
        print(f"  problem {k}: lambda_tsv {lt:.3g} alpha {info['alpha']:.3f} start gap (warm/cold) "
              f"{gaps[-1]:.3e} warm ITER {info['warm_iter']} cold ITER {info['cold_iter']} "
              f"accepted {info['warm_accepted']}", flush=True)
    print(f"  mean start gap ratio {np.mean(gaps):.3e}; mean ITER warm {np.mean(wi):.1f} cold {np.mean(ci):.1f}",
          flush=True)
    assert np.mean(gaps) < 0.5
