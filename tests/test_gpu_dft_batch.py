"""DFT mode with many right-hand sides on the FP64 matrix cores.

The reference multiplies A by one vector at a time (dgemv,
c/mfista_dft_lib.c:30-48).  mfista_dft_products computes R = y 1^T - A X and
G = A^T R for K vectors as GEMMs (v_mfma_f64_16x16x4f64), and
mfista_imaging_core_dft_batch runs up to 64 DFT-mode problems that share y and
A in lockstep on those products.  Checked against numpy (products), against
the single-problem GPU solver and the FP64 oracle (solves).  FP64 bars: image
relative L2 <= 1e-6, objective <= 1e-9 relative.
"""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [1, 5, 16, 37, 64])
def test_dft_products_match_numpy(K):
    rng = np.random.default_rng(100 + K)
    M, N = 301, 221  # not multiples of the 64-row tiles or the 16-wide inner chunks
    A = rng.standard_normal((M, N))
    X = rng.standard_normal((N, K))
    y = rng.standard_normal(M)
    print(f"case dft products M={M} N={N} K={K}", flush=True)
    R, G = sa.dft_products(A, X, y)
    Rr = y[:, None] - A @ X
    Gr = A.T @ Rr
    er = np.abs(R - Rr).max() / np.abs(Rr).max()
    eg = np.abs(G - Gr).max() / np.abs(Gr).max()
    print(f"  max rel error R {er:.2e} G {eg:.2e}", flush=True)
    assert er <= 1e-13 and eg <= 1e-13
    R0, G0 = sa.dft_products(A, X, None, adjoint=False)
    assert G0 is None and np.abs(R0 + A @ X).max() <= 1e-13 * np.abs(A @ X).max()


def _problem(seed, NX=16, NY=16, M=600, signed=False):
    rng = np.random.default_rng(seed)
    N = NX * NY
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    xs = rng.standard_normal(N) if signed else np.maximum(rng.standard_normal(N), 0)
    y = A @ xs + 0.05 * rng.standard_normal(M)
    return A, y


@pytest.mark.parametrize("case", ["nonneg", "signed_box"])
def test_dft_batch_matches_single_and_oracle(case):
    NX = NY = 16
    A, y = _problem(7, NX, NY, signed=(case != "nonneg"))
    nonneg = 1 if case == "nonneg" else 0
    box = None
    if case == "signed_box":
        box = np.ones(NX * NY, np.float32)
        box[::7] = 0.0
    l1s, tsvs = [0.02, 0.05, 0.1], [0.0, 0.1, 0.3, 1.0]
    grid = [(a, b) for a in l1s for b in tsvs]
    kw = dict(maxiter=3000, eps=1e-10)
    print(f"case dft batch {case}: {len(grid)} problems, M=600 N=256", flush=True)
    xs, rs, resid, info = sa.mfista_imaging_core_dft_batch(y, A, NX, NY, [g[0] for g in grid], [g[1] for g in grid],
                                                           cinit=1.0, nonneg_flag=nonneg, cl_box=box, **kw)
    print(f"  batch: {info['solve_ms']:.1f} ms, {info['total_trials']} trials", flush=True)
    worst_img = worst_obj = 0.0
    for k, (l1, ltsv) in enumerate(grid):
        x1, r1, _, res1 = sa.mfista_imaging_core_dft(y, A, NX, NY, lambda_l1=l1, lambda_tsv=ltsv, cinit=1.0,
                                                     nonneg_flag=nonneg, cl_box=box, **kw)
        di = np.linalg.norm(xs[k] - x1) / max(np.linalg.norm(x1), 1e-300)
        do = abs(rs[k]["finalcost"] - r1["finalcost"]) / r1["finalcost"]
        worst_img, worst_obj = max(worst_img, di), max(worst_obj, do)
        assert di <= 1e-6 and do <= 1e-9, (k, di, do)
        assert abs(rs[k]["ITER"] - r1["ITER"]) <= 0.02 * r1["ITER"] + 2, (k, rs[k]["ITER"], r1["ITER"])
        assert rs[k]["M"] == 600 and rs[k]["N"] == NX * NY and rs[k]["nonneg"] == nonneg
        assert np.abs(resid[k] - res1).max() <= 1e-6 * np.abs(res1).max()
        if box is not None:
            assert np.all(xs[k].ravel()[box == 0] == 0)
    print(f"  batch vs single: worst image relL2 {worst_img:.2e} objective rel {worst_obj:.2e}", flush=True)
    for k in (1, 10):
        l1, ltsv = grid[k]
        o = oracle.imaging_core_dft(y, A, NX, NY, kw["maxiter"], kw["eps"], l1, 0.0, ltsv, 1.0, nonneg=nonneg,
                                    cl_box=box)
        di = np.linalg.norm(xs[k] - o["x"]) / np.linalg.norm(o["x"])
        do = abs(rs[k]["finalcost"] - o["result"]["finalcost"]) / o["result"]["finalcost"]
        print(f"  problem {k} vs oracle: ITER {rs[k]['ITER']} / {o['result']['ITER']} image {di:.2e} "
              f"objective {do:.2e}", flush=True)
        assert di <= 1e-6 and do <= 1e-9
