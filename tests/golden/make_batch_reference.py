"""Writes tests/golden/batch_oracle.npz: the FP64 oracle's solutions of two
problems of each batched workload, at a fixed 3000 MFISTA iterations (no
stopping rule), 12-tap preset (c++/mfista_nufft_lib.cpp:680-927 restated in
oracle/mfista_oracle.c), nonneg, cinit = 1e4, xinit = 0:

  lambda grid  the 128^2 visibility set synth.make_nufft_problem(N=128,
               M=8000, seed=1004) with (lambda_l1, lambda_tsv) = (1, 100) and
               (10, 1000), two points of the 6 x 6 grid
               log10 lambda_l1 in linspace(-2, 3, 6), log10 lambda_tsv in
               linspace(0, 5, 6)
  movie        frames 3 and 20 of synth.make_movie_frames(nframes=32, N=128,
               seed=1004) (M between 5000 and 40000 per frame), lambda_l1 = 1,
               lambda_tsv = 100

Stored per problem: the image, the objective and the SHA-256 of the inputs.
Takes a few minutes on one core.  Run:  python tests/golden/make_batch_reference.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from sparseimaging_amd import synth  # noqa: E402

ITERS = 3000
GRID_POINTS = ((1.0, 100.0), (10.0, 1000.0))
MOVIE_FRAMES = (3, 20)


def sha(P):
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def solve(P, N, l1, ltsv):
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], N, N, ITERS, 0.0, l1, 0.0, ltsv,
                                  1e4, nonneg=1, preset=oracle.PRESET_CXX12, rel_tol=0.0)
    print(f"  ITER {o['result']['ITER']} finalcost {o['result']['finalcost']!r}", flush=True)
    return o["x"], o["result"]["finalcost"]


def main():
    out = {}
    P = synth.make_nufft_problem(N=128, M=8000, seed=1004)
    out["grid_sha256"] = np.array(sha(P))
    for k, (l1, ltsv) in enumerate(GRID_POINTS):
        print(f"lambda grid point {k}: lambda_l1={l1} lambda_tsv={ltsv}", flush=True)
        out[f"grid{k}_x"], F = solve(P, 128, l1, ltsv)
        out[f"grid{k}_F"] = np.array(F)
        out[f"grid{k}_lambda"] = np.array([l1, ltsv])
    frames = synth.make_movie_frames(nframes=32, N=128, seed=1004)
    for k, f in enumerate(MOVIE_FRAMES):
        print(f"movie frame {f}: M={frames[f]['M']}", flush=True)
        out[f"movie{k}_sha256"] = np.array(sha(frames[f]))
        out[f"movie{k}_frame"] = np.array(f)
        out[f"movie{k}_x"], F = solve(frames[f], 128, 1.0, 100.0)
        out[f"movie{k}_F"] = np.array(F)
    np.savez_compressed(os.path.join(HERE, "batch_oracle.npz"), **out)


if __name__ == "__main__":
    main()
