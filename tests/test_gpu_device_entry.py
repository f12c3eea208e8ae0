"""The device-pointer entry point (mfista_nufft_solve_device) is ordered after
the caller's work on the caller's stream.

The inputs are written by torch kernels queued just before the call, behind a
long-running kernel, with no synchronisation in between; the solve must read
the written values.  Reference for the solve itself: c/mfista_nufft_lib.c:909.
Compared with the host-pointer entry point at a fixed iteration count in the
deterministic mode (gridding without floating-point atomics), where two solves
of the same problem must agree bit for bit; a solve that read stale inputs
would be off by O(1).
"""
import numpy as np
import pytest

import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu

KW = dict(maxiter=300, eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4, nonneg_flag=1,
          preset=sa.PRESET_CXX12, deterministic=True)


def _host(P, N):
    x, r, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], N, N, **KW)
    return x.ravel(), r


@pytest.mark.parametrize("side_stream", [False, True])
def test_device_entry_orders_after_caller_stream(side_stream):
    import torch

    N = 64
    P = synth.make_nufft_problem(N=N, M=1500, seed=4242, amax=0.35)
    xh, rh = _host(P, N)
    dev = torch.device("cuda:0")
    print(f"case device entry, side_stream={side_stream}, N=64 M=1500, 300 iterations", flush=True)
    s = torch.cuda.Stream() if side_stream else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        t = {k: torch.zeros(P[k].size, dtype=torch.float64, device=dev) for k in ("u", "v", "vis_r", "vis_i", "std")}
        src = {k: torch.tensor(P[k], dtype=torch.float64).pin_memory() for k in t}
        big = torch.randn(4096, 4096, device=dev)
        for _ in range(8):  # keeps the stream busy for a while before the inputs are written
            big = big @ big
            big = big / big.norm()
        for k in t:
            t[k].copy_(src[k], non_blocking=True)
        x0 = torch.zeros(N * N, dtype=torch.float64, device=dev)
        xo = torch.full((N * N,), -1.0, dtype=torch.float64, device=dev)
        res, info = sa.solve_nufft_device(t["u"], t["v"], t["vis_r"], t["vis_i"], t["std"], N, N, x0, xo, **KW)
        xd = xo.cpu().numpy()
    l2 = np.linalg.norm(xd - xh) / np.linalg.norm(xh)
    dc = abs(res["finalcost"] - rh["finalcost"]) / rh["finalcost"]
    print(f"  device vs host entry: image relL2 {l2:.2e} objective rel {dc:.2e} engine {info['engine_used']}")
    assert info["engine_used"] == sa.ENGINE_PERSISTENT
    assert res["ITER"] == rh["ITER"] == 300
    assert np.array_equal(xd, xh) and res["finalcost"] == rh["finalcost"]


def test_deterministic_solves_repeat_bit_for_bit():
    """Deterministic mode: every gridding tile is summed by one workgroup in
    entry order, so two solves of a problem give the same bits (the default
    mode's FP32 atomics do not); both presets, L1+TSV and L1+TV."""
    P = synth.make_nufft_problem(N=96, M=4000, seed=77, amax=0.35)
    args = (P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 96, 96)
    for preset, ltv, ltsv in ((sa.PRESET_CXX12, 0.0, 30.0), (sa.PRESET_C24, 0.0, 30.0), (sa.PRESET_C24, 3.0, 0.0)):
        print(f"case deterministic repeat preset={preset} lambda_tv={ltv} lambda_tsv={ltsv} N=96 M=4000", flush=True)
        kw = dict(maxiter=400, eps=0.0, lambda_l1=1.0, lambda_tv=ltv, lambda_tsv=ltsv, cinit=1e4, nonneg_flag=1,
                  preset=preset, fgp_iter=20, deterministic=True)
        xa, ra, ia = sa.mfista_imaging_core_nufft(*args, **kw)
        xb, rb, ib = sa.mfista_imaging_core_nufft(*args, **kw)
        print(f"  ITER {ra['ITER']} {rb['ITER']} trials {ia['total_trials']} {ib['total_trials']} "
              f"max |dx| {np.abs(xa - xb).max():.3e}", flush=True)
        assert np.array_equal(xa, xb)
        assert ra["finalcost"] == rb["finalcost"] and ia["total_trials"] == ib["total_trials"]
