"""The Toeplitz loop of single NUFFT solves against the FP64 oracle and the
gridding loop.

The loop applies the data term's normal operator A^H W^2 A as a convolution
(FP64 FFTs on the padded grid, kernel and A^H W^2 V from a 24-tap FP64 NUFFT
adjoint at set-up) instead of gridding and degridding every iteration; the
reference (c/mfista_nufft_lib.c:557-787) evaluates the same operator with its
preset's NUFFT.  Parity bar as for every NUFFT solve: image relative L2
<= 1e-3 and final objective (the FP64 RESULT pass, with the preset's own
NUFFT) within 1e-6 of the oracle's.
"""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu


def _oracle(P, NX, NY, preset, maxiter, rel_tol, lambda_l1, lambda_tsv, nonneg=1, box=None):
    return oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], NX, NY, maxiter, 0.0,
                                     lambda_l1, 0.0, lambda_tsv, 1e4, nonneg=nonneg, cl_box=box, preset=preset,
                                     rel_tol=rel_tol)


def _gpu(P, NX, NY, preset, maxiter, rel_tol, lambda_l1, lambda_tsv, nonneg=1, box=None, loop=sa.LOOP_AUTO):
    return sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], NX, NY, maxiter=maxiter,
                                        eps=0.0, lambda_l1=lambda_l1, lambda_tsv=lambda_tsv, cinit=1e4,
                                        nonneg_flag=nonneg, cl_box=box, preset=preset, rel_tol=rel_tol, loop=loop)


def _compare(xg, rg, o):
    xo, ro = o["x"], o["result"]
    l2 = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    dc = abs(rg["finalcost"] - ro["finalcost"]) / abs(ro["finalcost"])
    return l2, dc


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_toeplitz_converged_matches_oracle(preset):
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    print(f"case toeplitz converged preset={preset} N=64 M=2000 rel_tol=1e-9", flush=True)
    o = _oracle(P, 64, 64, preset, 20000, 1e-9, 1.0, 10.0)
    xg, rg, info = _gpu(P, 64, 64, preset, 20000, 1e-9, 1.0, 10.0)
    l2, dc = _compare(xg, rg, o)
    print(f"  loop {info['loop_used']} ITER gpu={rg['ITER']} oracle={o['result']['ITER']} relL2={l2:.2e} "
          f"cost rel={dc:.2e} kernel {info['kernel_ms']:.2f} ms")
    assert info["loop_used"] == sa.LOOP_TOEPLITZ
    assert l2 <= 1e-3 and dc <= 1e-6


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_toeplitz_and_gridding_loops_at_fixed_iterations(preset):
    """Both loops and the oracle for the same 4000 iterations (past
    convergence) of a 96^2 problem with lambda_tsv = 100."""
    P = synth.make_nufft_problem(N=96, M=3000, seed=555, amax=0.35)
    print(f"case toeplitz vs gridding preset={preset} N=96 M=3000 4000 iterations", flush=True)
    o = _oracle(P, 96, 96, preset, 4000, 0.0, 1.0, 100.0)
    xt, rt, it = _gpu(P, 96, 96, preset, 4000, 0.0, 1.0, 100.0)
    xg, rg, ig = _gpu(P, 96, 96, preset, 4000, 0.0, 1.0, 100.0, loop=sa.LOOP_GRIDDING)
    lt, dt = _compare(xt, rt, o)
    lg, dg = _compare(xg, rg, o)
    print(f"  toeplitz: relL2={lt:.2e} cost rel={dt:.2e} ({it['kernel_ms']:.1f} ms); gridding: relL2={lg:.2e} "
          f"cost rel={dg:.2e} ({ig['kernel_ms']:.1f} ms)")
    assert it["loop_used"] == sa.LOOP_TOEPLITZ and ig["loop_used"] == sa.LOOP_GRIDDING
    assert rt["ITER"] == 4000 and rg["ITER"] == 4000
    assert lt <= 1e-3 and dt <= 1e-6
    assert lg <= 1e-3 and dg <= 1e-6


@pytest.mark.parametrize("N", [384, 512])
def test_toeplitz_large_images_match_gridding_loop(N):
    """Images wider than a workgroup (each thread loads two pixels of a row
    and forms two entries of a column in the pruned transforms; 384 = 3 x 128
    also takes the radix-3 stage): the Toeplitz loop, the persistent Toeplitz
    batch and the wide engine's Toeplitz steps against the gridding loop (whose
    stages do not share that code) for the same 8000 iterations, to the parity
    bars.  Both loops follow the oracle to those bars at 96^2 above; an oracle
    solve at this size takes minutes.  LOOP_AUTO picks the Toeplitz loop here
    too; LOOP_TOEPLITZ asks for it."""
    P = synth.make_nufft_problem(N=N, M=20000, seed=4242)
    print(f"case toeplitz large image N={N} M=20000 8000 iterations", flush=True)
    _, _, ia = _gpu(P, N, N, sa.PRESET_CXX12, 50, 0.0, 1.0, 100.0)
    assert ia["loop_used"] == sa.LOOP_TOEPLITZ
    xt, rt, it = _gpu(P, N, N, sa.PRESET_CXX12, 8000, 0.0, 1.0, 100.0, loop=sa.LOOP_TOEPLITZ)
    xg, rg, ig = _gpu(P, N, N, sa.PRESET_CXX12, 8000, 0.0, 1.0, 100.0, loop=sa.LOOP_GRIDDING)
    assert it["loop_used"] == sa.LOOP_TOEPLITZ and ig["loop_used"] == sa.LOOP_GRIDDING
    l2 = np.linalg.norm(xt - xg) / np.linalg.norm(xg)
    dc = abs(rt["finalcost"] - rg["finalcost"]) / rg["finalcost"]
    print(f"  toeplitz vs gridding: relL2={l2:.2e} cost rel={dc:.2e} ({it['kernel_ms']:.1f} / "
          f"{ig['kernel_ms']:.1f} ms)", flush=True)
    assert rt["ITER"] == 8000 and rg["ITER"] == 8000
    assert l2 <= 1e-3 and dc <= 1e-6
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=N, NY=N,
                maxiter=8000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, lambda_tsv=t) for t in (100.0, 300.0)]
    xp, rp, ip = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, loop=sa.LOOP_TOEPLITZ)
    xw, rw, iw = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    assert ip["loop_used"] == sa.LOOP_TOEPLITZ
    assert iw["engine_used"] == sa.ENGINE_STREAM and iw["loop_used"] == sa.LOOP_TOEPLITZ
    for name, xs, rs in (("persistent batch", xp, rp), ("wide", xw, rw)):
        l2 = np.linalg.norm(xs[0] - xt) / np.linalg.norm(xt)
        dc = abs(rs[0]["finalcost"] - rt["finalcost"]) / rt["finalcost"]
        l2b = np.linalg.norm(xs[1] - xp[1]) / np.linalg.norm(xp[1])
        print(f"  {name}: lambda_tsv 100 vs single relL2={l2:.2e} cost rel={dc:.2e}; lambda_tsv 300 vs the "
              f"persistent batch relL2={l2b:.2e}", flush=True)
        assert rs[0]["ITER"] == 8000 and l2 <= 1e-3 and dc <= 1e-6 and l2b <= 1e-3


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_toeplitz_signed_with_box(preset):
    """nonneg_flag = 0 (the signed soft threshold, c/mfista_tools.c:81-112)
    with a cl_box window, lambda_tsv = 30, 5000 iterations."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=77, amax=0.35)
    box = np.zeros((64, 64), dtype=np.float32)
    box[8:56, 10:54] = 1.0
    print(f"case toeplitz signed+box preset={preset} N=64 M=2000 5000 iterations", flush=True)
    o = _oracle(P, 64, 64, preset, 5000, 0.0, 1.0, 30.0, nonneg=0, box=box)
    xg, rg, info = _gpu(P, 64, 64, preset, 5000, 0.0, 1.0, 30.0, nonneg=0, box=box)
    l2, dc = _compare(xg, rg, o)
    print(f"  loop {info['loop_used']} relL2={l2:.2e} cost rel={dc:.2e} min x {xg.min():.3e}")
    assert info["loop_used"] == sa.LOOP_TOEPLITZ
    assert np.all(xg[box == 0] == 0.0)
    assert xg.min() < 0.0  # the signed threshold does produce negative pixels here
    assert l2 <= 1e-3 and dc <= 1e-6


def test_toeplitz_l1_only():
    """lambda_tsv = 0: the L1 solve (the image is the unique minimiser only up
    to the data term's null space, so the objective is the bar).  It converges
    slowly: at 3000 iterations the objective is still ~1e-6 from the optimum
    and two trajectories that differ in the last bits (the set-up adjoint's
    FP64 atomics) end up to 1.1e-6 apart, so both run 12000 iterations."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=99, amax=0.35)
    print("case toeplitz L1 only N=64 M=2000 12000 iterations", flush=True)
    o = _oracle(P, 64, 64, sa.PRESET_CXX12, 12000, 0.0, 5.0, 0.0)
    xg, rg, info = _gpu(P, 64, 64, sa.PRESET_CXX12, 12000, 0.0, 5.0, 0.0)
    l2, dc = _compare(xg, rg, o)
    print(f"  loop {info['loop_used']} relL2={l2:.2e} cost rel={dc:.2e}")
    assert info["loop_used"] == sa.LOOP_TOEPLITZ
    assert dc <= 1e-6


def test_loop_selection():
    """TV solves, deterministic mode, single-trial rounds and rectangular
    images keep the gridding loop (the presets size both axes' kernels from
    NX, so for NX != NY the reference's transform is not the exact one the
    Toeplitz loop applies), also when LOOP_TOEPLITZ asks for the other;
    LOOP_GRIDDING forces it."""
    P = synth.make_nufft_problem(N=32, M=600, seed=5, amax=0.3)
    base = dict(maxiter=50, eps=0.0, lambda_l1=1.0, cinit=1e4, nonneg_flag=1)
    args = (P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 32, 32)
    print("case loop selection N=32", flush=True)
    _, _, i0 = sa.mfista_imaging_core_nufft(*args, lambda_tsv=1.0, preset=sa.PRESET_CXX12, **base)
    _, _, i1 = sa.mfista_imaging_core_nufft(*args, lambda_tv=1.0, preset=sa.PRESET_C24, **base)
    _, _, i2 = sa.mfista_imaging_core_nufft(*args, lambda_tsv=1.0, preset=sa.PRESET_CXX12, deterministic=True, **base)
    _, _, i3 = sa.mfista_imaging_core_nufft(*args, lambda_tsv=1.0, preset=sa.PRESET_CXX12, single_trials=True, **base)
    _, _, i4 = sa.mfista_imaging_core_nufft(*args, lambda_tsv=1.0, preset=sa.PRESET_CXX12, loop=sa.LOOP_GRIDDING,
                                            **base)
    _, _, i5 = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 32, 48,
                                            lambda_tsv=1.0, preset=sa.PRESET_CXX12, **base)
    _, _, i6 = sa.mfista_imaging_core_nufft(*args, lambda_tv=1.0, preset=sa.PRESET_C24, loop=sa.LOOP_TOEPLITZ,
                                            **base)
    assert i0["loop_used"] == sa.LOOP_TOEPLITZ
    assert [i["loop_used"] for i in (i1, i2, i3, i4, i5, i6)] == [sa.LOOP_GRIDDING] * 6


def test_toeplitz_batch_matches_single_solves_and_oracle():
    """Three lambda_tsv values over one visibility set as one persistent batch
    (the Toeplitz loop, one set-up shared by the batch) and as single solves,
    4000 iterations: every problem against its single solve and the middle one
    against the oracle, to the parity bars."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=4000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, lambda_tsv=t) for t in (1.0, 10.0, 100.0)]
    print("case toeplitz batch of 3 (64^2, M=2000), 4000 iterations", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12)
    assert info["engine_used"] == sa.ENGINE_PERSISTENT and info["loop_used"] == sa.LOOP_TOEPLITZ
    for p, xb, rb in zip(probs, xs, rs):
        xg, rg, ig = _gpu(P, 64, 64, sa.PRESET_CXX12, 4000, 0.0, 1.0, p["lambda_tsv"])
        l2 = np.linalg.norm(xb - xg) / np.linalg.norm(xg)
        dc = abs(rb["finalcost"] - rg["finalcost"]) / rg["finalcost"]
        print(f"  lambda_tsv {p['lambda_tsv']:g}: ITER {rb['ITER']} batch vs single relL2={l2:.2e} cost rel={dc:.2e}")
        assert rb["ITER"] == 4000 and ig["loop_used"] == sa.LOOP_TOEPLITZ
        assert l2 <= 1e-3 and dc <= 1e-6
    o = _oracle(P, 64, 64, sa.PRESET_CXX12, 4000, 0.0, 1.0, 10.0)
    l2, dc = _compare(xs[1], rs[1], o)
    print(f"  lambda_tsv 10 vs oracle: relL2={l2:.2e} cost rel={dc:.2e}")
    assert l2 <= 1e-3 and dc <= 1e-6
    xg, rg, ig = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, loop=sa.LOOP_GRIDDING)
    assert ig["loop_used"] == sa.LOOP_GRIDDING


def test_toeplitz_wide_engine_matches_persistent_and_oracle():
    """The wide engine's Toeplitz steps (every stage one launch for all
    problems): three lambda_tsv values at a fixed 4000 iterations against the
    persistent Toeplitz batch and the oracle, to the parity bars."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=4000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, lambda_tsv=t) for t in (1.0, 10.0, 100.0)]
    print("case toeplitz wide engine, 3 problems (64^2, M=2000), 4000 iterations", flush=True)
    xw, rw, iw = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    assert iw["engine_used"] == sa.ENGINE_STREAM and iw["loop_used"] == sa.LOOP_TOEPLITZ
    xp, rp, ip = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12)
    assert ip["loop_used"] == sa.LOOP_TOEPLITZ
    for p, a, b, c, d in zip(probs, xw, rw, xp, rp):
        l2 = np.linalg.norm(a - c) / np.linalg.norm(c)
        dc = abs(b["finalcost"] - d["finalcost"]) / d["finalcost"]
        print(f"  lambda_tsv {p['lambda_tsv']:g}: wide vs persistent relL2={l2:.2e} cost rel={dc:.2e}")
        assert b["ITER"] == 4000 and l2 <= 1e-3 and dc <= 1e-6
    o = _oracle(P, 64, 64, sa.PRESET_CXX12, 4000, 0.0, 1.0, 10.0)
    l2, dc = _compare(xw[1], rw[1], o)
    print(f"  wide lambda_tsv 10 vs oracle: relL2={l2:.2e} cost rel={dc:.2e}")
    assert l2 <= 1e-3 and dc <= 1e-6


def test_toeplitz_wide_target_objective_stop():
    """Target objectives on the wide engine's Toeplitz steps: run 600
    iterations, then the same problems with those objectives as targets.  The
    steps compare the exact operator's objective with the target, which is
    the 12-tap preset's objective (the RESULT pass's): the two differ by about
    1e-8 (relative) near a solution, so each problem stops within a few
    iterations of 600, at its target to within 1e-7."""
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, maxiter=600, lambda_tsv=t) for t in (1.0, 10.0, 100.0)]
    print("case toeplitz wide engine target objective: 600 iterations, then the same objectives as targets",
          flush=True)
    _, r0, i0 = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    assert i0["loop_used"] == sa.LOOP_TOEPLITZ
    tprobs = [dict(p, maxiter=20000, target_objective=r["finalcost"]) for p, r in zip(probs, r0)]
    _, r1, info = sa.mfista_imaging_core_nufft_batch(tprobs, preset=sa.PRESET_CXX12)
    assert info["engine_used"] == sa.ENGINE_STREAM and info["loop_used"] == sa.LOOP_TOEPLITZ
    for a, b in zip(r0, r1):
        print(f"  target {a['finalcost']:.10e}: ITER {b['ITER']} objective {b['finalcost']:.10e}", flush=True)
        assert abs(b["ITER"] - 600) <= 20
        assert b["finalcost"] <= a["finalcost"] * (1 + 1e-7)
