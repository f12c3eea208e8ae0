"""Writes tests/golden/cfg2_oracle.npz: the FP64 oracle's solutions of the
256^2 NUFFT L1+TSV workload (the benchmark problem).

Workload: synth.make_nufft_problem(N=256, M=20000, seed=1002) -- ring of
radius 40 px plus 3 Gaussians, 12 elliptical uv arcs, 1% noise -- solved with
the C++ 12-tap preset (c++/mfista_nufft_lib.cpp:680-927 restated in
oracle/mfista_oracle.c), lambda_l1 = 1, lambda_tsv = 100, nonneg, cinit = 1e4,
xinit = 0.  Three end points:
  rule   stop when cost[t-50] - cost[t] <= 1e-6 cost[t], t >= 100 (the
         benchmark's rule, the relative form of c/mfista_nufft_lib.c:741)
  4000   a fixed 4000 iterations
  20000  a fixed 20000 iterations (converged: the objective changes by less
         than 1e-9 relative over the last 4000)
Stored: the three images, objectives and iteration counts, the cost history of
the rule run, and the SHA-256 of the input arrays (so a test can check that
synth still generates the same problem).  Takes about 20 minutes on one core.
Run:  python tests/golden/make_cfg2_reference.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from sparseimaging_amd import synth  # noqa: E402

N, M, SEED = 256, 20000, 1002
LAMBDA_L1, LAMBDA_TSV, CINIT = 1.0, 100.0, 1e4


def inputs_sha256(P):
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def solve(P, maxiter, rel_tol, trace=False):
    return oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], N, N, maxiter, 0.0,
                                     LAMBDA_L1, 0.0, LAMBDA_TSV, CINIT, nonneg=1, preset=oracle.PRESET_CXX12,
                                     rel_tol=rel_tol, trace=trace)


def main():
    P = synth.make_nufft_problem(N=N, M=M, seed=SEED)
    out = {"sha256": np.array(inputs_sha256(P))}
    for name, maxiter, tol in (("rule", 50000, 1e-6), ("4000", 4000, 0.0), ("20000", 20000, 0.0)):
        o = solve(P, maxiter, tol, trace=(name == "rule"))
        r = o["result"]
        out[f"x_{name}"] = o["x"]
        out[f"F_{name}"] = np.array(r["finalcost"])
        out[f"ITER_{name}"] = np.array(r["ITER"])
        if name == "rule":
            out["cost_rule"] = o["cost"]
        print(f"{name}: ITER {r['ITER']} finalcost {r['finalcost']!r}", flush=True)
    np.savez_compressed(os.path.join(HERE, "cfg2_oracle.npz"), **out)


if __name__ == "__main__":
    main()
