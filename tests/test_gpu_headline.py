"""Parity at the benchmark workload: 256^2 NUFFT L1+TSV, M = 20000 on 12
elliptical arcs (seed 1002), 12-tap preset, lambda_l1 = 1, lambda_tsv = 100,
nonneg, cinit = 1e4 -- the problem bench.py times.

Checked against the FP64 oracle's solutions in tests/golden/cfg2_oracle.npz
(written by tests/golden/make_cfg2_reference.py; the loop restated is
c/mfista_nufft_lib.c:557-787 with the C++ preset of
c++/mfista_nufft_lib.cpp:85-184).  Parity bar: image relative L2 <= 1e-3 and
objective relative error <= 1e-6, at three end points:
  * the benchmark's rule, cost[t-50] - cost[t] <= 1e-6 cost[t];
  * 4000 iterations;
  * 20000 iterations (converged).
At the rule's stop both solvers are still about 4e-5 above the converged
objective: the rule measures progress, not distance to the optimum, and near
the stop the iterate still moves by about 3e-4 (relative L2) per iteration.
The GPU and the oracle can therefore stop a few iterations apart (575 against
580 seen), and their images then differ by that motion, not by arithmetic:
the oracle's own iterates 575 and 580 differ by 1.5e-3.  So at the rule the
image bar is applied to the oracle's iterate at the GPU's stopping count (the
serial oracle, run here for that many iterations), and the objective bar to
both that iterate and the oracle's own rule stop.  The image distance between
the two rule stops is printed and held to 1e-2 (about 30 iterations apart).
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
REF = np.load(os.path.join(HERE, "golden", "cfg2_oracle.npz"))
KW = dict(eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=100.0, cinit=1e4, nonneg_flag=1,
          preset=sa.PRESET_CXX12)


@pytest.fixture(scope="module")
def problem():
    P = synth.make_nufft_problem(N=256, M=20000, seed=1002)
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    assert h.hexdigest() == str(REF["sha256"]), "synth no longer generates the benchmark problem"
    return P


@pytest.mark.parametrize("end", ["rule", "4000", "20000"])
def test_headline_parity(problem, end):
    P = problem
    maxiter, tol = (50000, 1e-6) if end == "rule" else (int(end), 0.0)
    print(f"case headline 256^2 M=20000 12 taps lambda_tsv=100, end point {end}", flush=True)
    x, r, info = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 256, 256,
                                              maxiter=maxiter, rel_tol=tol, **KW)
    xo, Fo = REF[f"x_{end}"], float(REF[f"F_{end}"])
    l2 = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    dF = abs(r["finalcost"] - Fo) / Fo
    gap = (r["finalcost"] - float(REF["F_20000"])) / float(REF["F_20000"])
    print(f"  ITER gpu {r['ITER']} oracle {int(REF[f'ITER_{end}'])}; image relL2 {l2:.3e}; objective rel "
          f"{dF:.3e} (gpu {r['finalcost']:.10e} oracle {Fo:.10e}); gap to the converged objective {gap:.2e}; "
          f"{info['kernel_ms']:.1f} ms", flush=True)
    assert dF <= 1e-6
    if end != "rule":
        assert l2 <= 1e-3
        return
    assert l2 <= 1e-2
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 256, 256, r["ITER"], 0.0,
                                  KW["lambda_l1"], 0.0, KW["lambda_tsv"], KW["cinit"], nonneg=1,
                                  preset=oracle.PRESET_CXX12, rel_tol=0.0)
    xs, Fs = o["x"], o["result"]["finalcost"]
    l2s = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
    dFs = abs(r["finalcost"] - Fs) / Fs
    print(f"  oracle at the GPU's {r['ITER']} iterations: image relL2 {l2s:.3e}; objective rel {dFs:.3e}",
          flush=True)
    assert l2s <= 1e-3
    assert dFs <= 1e-6
