"""The batched and learned workloads at their full sizes:

  * 6 x 6 lambda grid (log10 lambda_l1 in [-2, 3], log10 lambda_tsv in [0, 5])
    over one 128^2 visibility set, M = 8000 on 12 arcs, in one batch launch;
  * 32 movie frames of 128^2 with M from 5000 to 40000 per frame, in one batch
    launch;
  * the learned warm start with the shipped 256^2 network on held-out
    problems, with the acceptance rule.

Every batched problem is compared with the same problem solved on its own
(both in the deterministic mode, a fixed 3000 iterations), and two problems
of each batch with the FP64 oracle's solutions at the same iteration count
(tests/golden/batch_oracle.npz, written by make_batch_reference.py).  There is
no reference counterpart to a batch (a batch is a loop of
c/mfista_nufft_lib.c:909 calls).  Bars: image relative L2 <= 1e-3 and
objective <= 1e-6 relative.
"""
import hashlib
import os

import numpy as np
import pytest

import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REF = np.load(os.path.join(HERE, "golden", "batch_oracle.npz"))
ITERS = 3000


def _sha(P):
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def _rel(x, y):
    ny = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / ny) if ny > 0 else float(np.linalg.norm(x))


def _check_batch(tag, probs, xs, rs):
    worst_img = worst_obj = 0.0
    for k, (p, xb, rb) in enumerate(zip(probs, xs, rs)):
        x1, r1, _ = sa.mfista_imaging_core_nufft(p["u_dx"], p["v_dy"], p["vis_r"], p["vis_i"], p["vis_std"],
                                                 p["NX"], p["NY"], maxiter=ITERS, eps=0.0, lambda_l1=p["lambda_l1"],
                                                 lambda_tsv=p["lambda_tsv"], cinit=1e4, nonneg_flag=1,
                                                 preset=sa.PRESET_CXX12, deterministic=True)
        di = _rel(xb, x1)
        do = abs(rb["finalcost"] - r1["finalcost"]) / abs(r1["finalcost"])
        worst_img, worst_obj = max(worst_img, di), max(worst_obj, do)
        assert rb["ITER"] == r1["ITER"], (tag, k, rb["ITER"], r1["ITER"])
        assert di <= 1e-3 and do <= 1e-6, (tag, k, di, do)
    print(f"  {tag}: {len(probs)} problems, batch vs single: worst image relL2 {worst_img:.2e}, "
          f"worst objective rel {worst_obj:.2e}", flush=True)


def _check_oracle(tag, xb, rb, xo, Fo):
    di = _rel(xb, xo)
    do = abs(rb["finalcost"] - Fo) / abs(Fo)
    print(f"  {tag} vs oracle: image relL2 {di:.2e} objective rel {do:.2e}", flush=True)
    assert di <= 1e-3 and do <= 1e-6


def test_lambda_grid_6x6_128_batch():
    P = synth.make_nufft_problem(N=128, M=8000, seed=1004)
    assert _sha(P) == str(REF["grid_sha256"])
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=128, NY=128,
                maxiter=ITERS, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0)
    probs = [dict(base, lambda_l1=10.0 ** a, lambda_tsv=10.0 ** b)
             for a in np.linspace(-2, 3, 6) for b in np.linspace(0, 5, 6)]
    print("case lambda grid 6x6, 128^2, M=8000, 3000 iterations, one batch launch", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, deterministic=True)
    print(f"  batch kernel {info['kernel_ms']:.1f} ms", flush=True)
    _check_batch("lambda grid", probs, xs, rs)
    for k in range(2):
        l1, ltsv = REF[f"grid{k}_lambda"]
        j = [i for i, p in enumerate(probs) if p["lambda_l1"] == l1 and p["lambda_tsv"] == ltsv][0]
        _check_oracle(f"grid point ({l1:g}, {ltsv:g})", xs[j], rs[j], REF[f"grid{k}_x"], float(REF[f"grid{k}_F"]))


def test_movie_32_frames_128_batch():
    frames = synth.make_movie_frames(nframes=32, N=128, seed=1004)
    probs = [dict(u_dx=f["u"], v_dy=f["v"], vis_r=f["vis_r"], vis_i=f["vis_i"], vis_std=f["std"], NX=128, NY=128,
                  maxiter=ITERS, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_l1=1.0, lambda_tsv=100.0, lambda_tv=0.0)
             for f in frames]
    Ms = [f["M"] for f in frames]
    print(f"case movie 32 frames 128^2, M {min(Ms)}..{max(Ms)}, 3000 iterations, one batch launch", flush=True)
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, deterministic=True)
    print(f"  batch kernel {info['kernel_ms']:.1f} ms", flush=True)
    _check_batch("movie", probs, xs, rs)
    for k in range(2):
        f = int(REF[f"movie{k}_frame"])
        assert _sha(frames[f]) == str(REF[f"movie{k}_sha256"])
        _check_oracle(f"frame {f}", xs[f], rs[f], REF[f"movie{k}_x"], float(REF[f"movie{k}_F"]))


def test_learned_warm_start_256_acceptance_rule():
    """The shipped network (4 levels, 32 base channels, 256^2) on two held-out
    problems: the solve with verify=True returns the lower of the warm and
    cold objectives, and an accepted warm start reaches the cold objective to
    1e-6 (both run to a 1e-9 relative tolerance)."""
    from sparseimaging_amd import learn

    net = learn.UNet.load(learn.WEIGHTS)
    assert net.meta.get("N") == 256
    rng = np.random.default_rng(9100)  # training used seeds 5000-7999, the bench 9000
    kw = dict(maxiter=50000, eps=0.0, cinit=1e4, nonneg_flag=1, rel_tol=1e-9)
    for k in range(2):
        P = learn.make_problem(256, 20000, rng)
        print(f"case learned warm start, held-out problem {k}, 256^2 M=20000", flush=True)
        x, r, info = learn.solve_warm(net, P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 256, lambda_l1=1.0,
                                      lambda_tsv=100.0, verify=True, **kw)
        cold = info["cold_finalcost"]
        print(f"  warm accepted {info['warm_accepted']}: warm ITER {info['warm_iter']} cold ITER "
              f"{info['cold_iter']}; returned objective {r['finalcost']:.10e} cold {cold:.10e}", flush=True)
        assert r["finalcost"] <= cold
        if info["warm_accepted"]:
            assert (cold - r["finalcost"]) / cold <= 1e-6
