"""The wide engine (mfista_imaging_core_nufft_batch with engine=ENGINE_STREAM):
the stream engine's loop (c/mfista_nufft_lib.c:557-787) with every kernel
launched once for all problems of a batch and one step replayed as a hipGraph.
There is no reference counterpart to a batch (a batch is a loop of
c/mfista_nufft_lib.c:909 calls).

* lambda grids over one visibility set, both presets, L1-only and L1+TSV
  problems mixed, a fixed 4000 iterations, every problem against the FP64
  oracle run for as many;
* problems on different visibility sets (uneven M), one with nonneg_flag = 0,
  one with a CLEAN box, against the oracle (4000 iterations);
* the full-size batched workloads (the 6x6 lambda grid at 128^2, M = 8000,
  and 32 movie frames at 128^2, M 5000-40000), a fixed 3000 iterations, two
  problems of each against the oracle's solutions at the same iteration count
  (tests/golden/batch_oracle.npz).
Bars: image relative L2 <= 1e-3, objective <= 1e-6 relative.  The small
problems run a fixed iteration count because several of them converge slowly
(a 1e-9 relative rule stops the two solvers at points whose images differ by
~3e-3 while their objectives agree to 1e-7).  The L1-only problem (M = 2000
visibilities, 4096 unknowns) is not strictly convex: only its objective is
held to the bar, its image is printed.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REF = np.load(os.path.join(HERE, "golden", "batch_oracle.npz"))


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def _oracle(p, preset):
    o = oracle.imaging_core_nufft(p["u_dx"], p["v_dy"], p["vis_r"], p["vis_i"], p["vis_std"], p["NX"], p["NY"],
                                  p["maxiter"], 0.0, p["lambda_l1"], 0.0, p["lambda_tsv"], p["cinit"],
                                  nonneg=p["nonneg_flag"], cl_box=p.get("cl_box"), preset=preset, rel_tol=0.0)
    return o["x"], o["result"]["finalcost"]


def _check(tag, x, r, xo, Fo, image=True):
    di = _rel(x, xo)
    do = abs(r["finalcost"] - Fo) / abs(Fo)
    print(f"  {tag}: ITER {r['ITER']} image relL2 {di:.2e} objective rel {do:.2e}", flush=True)
    assert do <= 1e-6, (tag, di, do)
    if image:
        assert di <= 1e-3, (tag, di, do)


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_wide_lambda_grid_matches_oracle(preset):
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=4000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, lambda_tsv=t) for t in (0.0, 1.0, 10.0, 100.0)]
    print(f"case wide engine lambda grid x4, 64^2 M=2000 preset {preset}, 4000 iterations", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=preset, engine=sa.ENGINE_STREAM)
    assert info["engine_used"] == sa.ENGINE_STREAM
    print(f"  wide batch {info['kernel_ms']:.1f} ms", flush=True)
    for p, x, r in zip(probs, xs, rs):
        assert r["ITER"] == 4000
        xo, Fo = _oracle(p, preset)
        _check(f"lambda_tsv {p['lambda_tsv']:g}", x, r, xo.reshape(64, 64), Fo, image=p["lambda_tsv"] > 0)


def test_wide_uneven_sets_signed_and_box():
    sets = [synth.make_nufft_problem(N=64, M=M, seed=2000 + M, amax=0.35) for M in (1500, 2500, 3500)]
    box = np.ones((64, 64), dtype=np.float32)
    box[:, :16] = 0
    probs = []
    for k, P in enumerate(sets):
        p = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                 maxiter=4000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0, lambda_tsv=10.0)
        if k == 0:
            p["nonneg_flag"] = 0
        if k == 1:
            p["cl_box"] = box
        probs.append(p)
    print("case wide engine, 3 visibility sets (M 1500/2500/3500), signed L1, CLEAN box, 4000 iterations",
          flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    assert info["engine_used"] == sa.ENGINE_STREAM
    for k, (p, x, r) in enumerate(zip(probs, xs, rs)):
        xo, Fo = _oracle(p, sa.PRESET_CXX12)
        _check(f"set {k}", x, r, xo.reshape(64, 64), Fo)
    assert np.all(xs[1][:, :16] == 0)


def _sha(P):
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def test_wide_lambda_grid_6x6_128():
    P = synth.make_nufft_problem(N=128, M=8000, seed=1004)
    assert _sha(P) == str(REF["grid_sha256"])
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=128, NY=128,
                maxiter=3000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0)
    probs = [dict(base, lambda_l1=10.0 ** a, lambda_tsv=10.0 ** b)
             for a in np.linspace(-2, 3, 6) for b in np.linspace(0, 5, 6)]
    print("case wide engine lambda grid 6x6, 128^2, M=8000, 3000 iterations", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    print(f"  wide batch {info['kernel_ms']:.1f} ms", flush=True)
    assert all(r["ITER"] == 3000 for r in rs)
    for k in range(2):
        l1, ltsv = REF[f"grid{k}_lambda"]
        j = [i for i, p in enumerate(probs) if p["lambda_l1"] == l1 and p["lambda_tsv"] == ltsv][0]
        _check(f"grid point ({l1:g}, {ltsv:g}) vs oracle", xs[j], rs[j], REF[f"grid{k}_x"], float(REF[f"grid{k}_F"]))


def test_wide_movie_32_frames_128():
    frames = synth.make_movie_frames(nframes=32, N=128, seed=1004)
    probs = [dict(u_dx=f["u"], v_dy=f["v"], vis_r=f["vis_r"], vis_i=f["vis_i"], vis_std=f["std"], NX=128, NY=128,
                  maxiter=3000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_l1=1.0, lambda_tsv=100.0, lambda_tv=0.0)
             for f in frames]
    print("case wide engine movie 32 frames 128^2, M 5000..40000, 3000 iterations", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    print(f"  wide batch {info['kernel_ms']:.1f} ms", flush=True)
    for k in range(2):
        f = int(REF[f"movie{k}_frame"])
        assert _sha(frames[f]) == str(REF[f"movie{k}_sha256"])
        _check(f"frame {f} vs oracle", xs[f], rs[f], REF[f"movie{k}_x"], float(REF[f"movie{k}_F"]))


def test_wide_compaction_uneven_iteration_counts():
    """Problems that finish at different steps: the wide engine re-captures
    its step for the problems still running each time they have halved, so
    the late problems run on smaller launches (here 8, then 4, then 2).
    Eight problems with maxiter 1000 .. 4000 (no stopping rule), every one
    against the oracle run for its own count.  (Earlier stops are not used:
    at 250 iterations the objective still falls fast enough that rounding in
    the trajectory alone moves it by ~1e-6.)"""
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    its = [1000, 1000, 1500, 1500, 2000, 2000, 3000, 4000]
    probs = [dict(base, maxiter=m, lambda_tsv=10.0 + k) for k, m in enumerate(its)]
    print("case wide engine compaction, 8 problems, maxiter 1000..4000", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM)
    assert info["engine_used"] == sa.ENGINE_STREAM
    for p, x, r in zip(probs, xs, rs):
        assert r["ITER"] == p["maxiter"]
        xo, Fo = _oracle(p, sa.PRESET_CXX12)
        _check(f"maxiter {p['maxiter']}", x, r, xo.reshape(64, 64), Fo)


def test_wide_target_objective_stop():
    """A target objective stops a problem at the first iteration whose
    objective is at or below it: run 600 iterations, then the same problems
    with that objective as the target (no stopping rule, maxiter 20000); each
    stops within a few iterations of 600 at or below its target.  A batch with
    targets runs on the wide engine even when the persistent one is asked for.
    (Gridding steps: their objective is the preset's own, as the RESULT
    pass's; the Toeplitz steps' target test is in test_gpu_toeplitz.py.)"""
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, maxiter=600, lambda_tsv=t) for t in (1.0, 10.0, 100.0)]
    print("case wide engine target objective: 600 iterations, then the same objectives as targets", flush=True)
    _, r0, _ = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, engine=sa.ENGINE_STREAM,
                                                  loop=sa.LOOP_GRIDDING)
    tprobs = [dict(p, maxiter=20000, target_objective=r["finalcost"]) for p, r in zip(probs, r0)]
    _, r1, info = sa.mfista_imaging_core_nufft_batch(tprobs, preset=sa.PRESET_CXX12, loop=sa.LOOP_GRIDDING)
    assert info["engine_used"] == sa.ENGINE_STREAM and info["loop_used"] == sa.LOOP_GRIDDING
    for a, b in zip(r0, r1):
        print(f"  target {a['finalcost']:.10e}: ITER {b['ITER']} objective {b['finalcost']:.10e}", flush=True)
        assert abs(b["ITER"] - 600) <= 5
        assert b["finalcost"] <= a["finalcost"] * (1 + 1e-9)
