"""Pins the optimum fixture (ref_nufft_solution.npz) that the GPU
reference-data test compares against, without re-running the solve.

The fixture holds the optimum of c/nufft_data.txt with the author's
parameters (lambda_l1 = lambda_tsv = 100, nonneg), found by L-BFGS-B on the
oracle's FP64 operators (make_golden.py).  Checks that do not go through the
oracle's MFISTA loop:
  * its stored finalcost is the objective of its stored image, recomputed
    with the FP64 result pass (c/mfista_nufft_lib.c:791-905);
  * the image satisfies the KKT conditions of the nonneg L1+TSV problem:
    with g = -Re A^H W^2 (V - A x) + lambda_l1 + lambda_tsv grad TSV(x),
    g = 0 on the support and g >= 0 off it, to 1e-3 of the data-term
    gradient's scale (a dropped term or a wrong sign misses by O(1));
  * it is at least as good as the oracle's own MFISTA run at a 1e-8
    tolerance, whose objective the fixture also records."""
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("preset", [oracle.PRESET_CXX12, oracle.PRESET_C24])
def test_reference_solution_fixture(preset):
    d = np.load(os.path.join(G, "ref_nufft_data.npz"))
    s = np.load(os.path.join(G, "ref_nufft_solution.npz"))
    N = int(d["NX"])
    l1 = ltsv = 100.0
    x2 = s[f"x{preset}"].astype(np.float64)
    x = x2.ravel()
    r = oracle.result_nufft(d["u"], d["v"], d["vis_r"], d["vis_i"], d["std"], x2, l1, 0.0, ltsv, preset=preset)
    fo = float(s[f"finalcost{preset}"])
    assert abs(r["finalcost"] - fo) <= 1e-9 * fo

    V = d["vis_r"] + 1j * d["vis_i"]
    w2 = 1.0 / d["std"] ** 2
    Ax = oracle.nufft_forward(d["u"], d["v"], x2, preset=preset)
    adj = oracle.nufft_adjoint(d["u"], d["v"], w2 * (V - Ax), N, N, preset=preset).ravel()
    g = -adj + l1 + ltsv * oracle.d_tsv_flat(x, N, N)
    scale = np.abs(adj).max()
    on = x > 0
    assert on.sum() > 100
    assert np.abs(g[on]).max() <= 1e-4 * scale
    assert g[~on].min() >= -1e-4 * scale
    assert fo <= float(s[f"mfista_finalcost{preset}"])
