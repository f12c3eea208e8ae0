"""Batched NUFFT solves (mfista_imaging_core_nufft_batch): many independent
problems in one persistent launch.

Each problem of a batch runs the same fixed number of MFISTA iterations as the
single-solve entry point and the FP64 oracle (no stopping rule: on these small
problems the objective is flat enough that two correct runs stopped by a 1e-9
relative rule differ by up to ~2e-3 in the image), and must agree with both to
image relative L2 <= 1e-3 and objective <= 1e-6 relative.  Convergence to the
optimum is tested on single solves (test_gpu_solve_nufft.py)."""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu


def _close(tag, xb, rb, xs, rs, img_tol=1e-3, cost_tol=1e-6):
    l2 = np.linalg.norm(xb - xs) / max(np.linalg.norm(xs), 1e-300)
    dc = abs(rb["finalcost"] - rs["finalcost"]) / abs(rs["finalcost"])
    print(f"  {tag}: ITER batch={rb['ITER']} single={rs['ITER']} relL2={l2:.2e} cost rel={dc:.2e}", flush=True)
    assert l2 <= img_tol, (tag, l2)
    assert dc <= cost_tol, (tag, dc)


def test_lambda_grid_one_visibility_set():
    """A 3 x 3 (lambda_l1, lambda_tsv) grid over one visibility set (shared set-up)."""
    P = synth.make_nufft_problem(N=64, M=1500, seed=41, amax=0.35)
    grid = [(l1, lt) for l1 in (0.3, 1.0, 3.0) for lt in (3.0, 10.0, 30.0)]
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=1500, eps=0.0, cinit=1e4, nonneg_flag=1)
    probs = [dict(base, lambda_l1=l1, lambda_tsv=lt) for l1, lt in grid]
    print(f"case lambda grid 3x3, 64x64, M=1500, CXX-12", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12)
    print(f"  batch: {info['solve_ms']:.1f} ms, kernel {info['kernel_ms']:.1f} ms", flush=True)
    for k, (l1, lt) in enumerate(grid):
        x1, r1, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                                 maxiter=1500, eps=0.0, lambda_l1=l1, lambda_tsv=lt, cinit=1e4,
                                                 nonneg_flag=1, preset=sa.PRESET_CXX12)
        _close(f"l1={l1} tsv={lt}", xs[k], rs[k], x1, r1)
        assert rs[k]["M"] == 1500 and rs[k]["ITER"] == r1["ITER"] == 1500
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, 1500, 0.0, 1.0, 0.0,
                                  10.0, 1e4, nonneg=1, preset=oracle.PRESET_CXX12)
    _close("oracle l1=1 tsv=10", xs[4], rs[4], o["x"], o["result"])


def test_movie_frames_uneven_visibility_counts():
    """Frames with their own (u,v) sets and uneven M (separate set-ups), C-24,
    fewer problems than XCDs."""
    probs, frames = [], []
    for f, M in enumerate((900, 1600, 2500)):
        P = synth.make_nufft_problem(N=64, M=M, seed=60 + f, amax=0.35)
        frames.append(P)
        probs.append(dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"],
                          NX=64, NY=64, maxiter=1500, eps=0.0, lambda_l1=1.0, lambda_tsv=10.0, cinit=1e4,
                          nonneg_flag=1))
    print("case movie 3 frames, 64x64, M = 900/1600/2500, C-24", flush=True)
    xs, rs, _ = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_C24)
    for f, P in enumerate(frames):
        x1, r1, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                                 maxiter=1500, eps=0.0, lambda_l1=1.0, lambda_tsv=10.0,
                                                 cinit=1e4, nonneg_flag=1, preset=sa.PRESET_C24)
        _close(f"frame {f}", xs[f], rs[f], x1, r1)
        assert rs[f]["M"] == P["u"].size // 2 and rs[f]["ITER"] == 1500
    P = frames[2]
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, 1500, 0.0, 1.0, 0.0,
                                  10.0, 1e4, nonneg=1, preset=oracle.PRESET_C24)
    _close("oracle frame 2", xs[2], rs[2], o["x"], o["result"])


def test_batch_tv_and_invalid_nonneg():
    """L1+TV problems (FGP in the batch kernel) next to a problem whose
    nonneg_flag is invalid (ITER = 0, c/mfista_nufft_lib.c:634-637)."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=77, amax=0.35, kind="tv", coverage="disc")
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=300, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tsv=0.0)
    probs = [dict(base, lambda_l1=0.5, lambda_tv=5.0), dict(base, lambda_l1=1.0, lambda_tv=2.0),
             dict(base, lambda_l1=1.0, lambda_tsv=10.0, nonneg_flag=2)]
    print("case TV batch + invalid nonneg_flag, 64x64, C-24, fgp=30", flush=True)
    xs, rs, _ = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_C24, fgp_iter=30)
    for k in range(2):
        x1, r1, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                                 maxiter=300, eps=0.0, lambda_l1=probs[k]["lambda_l1"],
                                                 lambda_tv=probs[k]["lambda_tv"], cinit=1e4, nonneg_flag=1,
                                                 preset=sa.PRESET_C24, fgp_iter=30)
        _close(f"tv problem {k}", xs[k], rs[k], x1, r1)
    assert rs[2]["ITER"] == 0


def test_batch_rejects_both_regularisers():
    P = synth.make_nufft_problem(N=32, M=300, seed=3, amax=0.3)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=32, NY=32)
    with pytest.raises(sa.MfistaError) as e:
        sa.mfista_imaging_core_nufft_batch([dict(base, lambda_tv=1.0, lambda_tsv=1.0)])
    assert e.value.status == -1
