"""Pins the oracle's signed-L1 and boxed paths against an independent
optimiser (CPU only).

nonneg_flag = 0 uses the signed soft threshold (c/mfista_tools.c:81-112) in
the L1+TSV loops and FGP_L1 (c/mfista_TV_lib.c:124-253) in the L1+TV loops.
Splitting x = p - q with p, q >= 0 turns lambda |x|_1 into the linear
lambda sum(p + q), so the problem is smooth and bound-constrained and scipy's
L-BFGS-B solves it independently of the oracle (at the optimum p q = 0).  A
CLEAN box (cl_box == 0) fixes those pixels at 0: bounds (0, 0).
"""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle




def _tsv_fg(x, NX, NY):
    return oracle.tsv_flat(x, NX, NY), oracle.d_tsv_flat(x, NX, NY)


def _solve_split(smooth_fg, n, l1, fixed=None):
    bounds = [(0, None)] * (2 * n)
    if fixed is not None:
        for k in np.flatnonzero(fixed):
            bounds[k] = (0, 0)
            bounds[n + k] = (0, 0)

    def fg(pq):
        x = pq[:n] - pq[n:]
        f, g = smooth_fg(x)
        return f + l1 * pq.sum(), np.concatenate([g + l1, -g + l1])

    r = minimize(fg, np.zeros(2 * n), jac=True, method="L-BFGS-B", bounds=bounds,
                 options={"maxiter": 100000, "maxfun": 200000, "ftol": 1e-16, "gtol": 1e-11})
    return r.x[:n] - r.x[n:], r.fun


def _dft_problem(seed, NX=8, NY=8, M=120):
    rng = np.random.default_rng(seed)
    N = NX * NY
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    xs = np.where(rng.uniform(size=N) < 0.5, rng.standard_normal(N), 0.0)
    y = A @ xs + 0.05 * rng.standard_normal(M)
    return A, y


def test_dft_signed_l1_tsv_reaches_optimum():
    """The MFISTA loop with the signed soft threshold (shared by the NUFFT,
    FFT and DFT cores; restated once in the oracle's mfista_loop), on a
    DFT-mode problem whose optimum L-BFGS-B finds quickly."""
    NX = NY = 8
    A, y = _dft_problem(61, NX, NY)
    l1, ltsv = 0.05, 0.2

    def smooth(x):
        r = y - A @ x
        t, dt = _tsv_fg(x, NX, NY)
        return 0.5 * r @ r + ltsv * t, -A.T @ r + ltsv * dt

    xb, fb = _solve_split(smooth, NX * NY, l1)
    o = oracle.imaging_core_dft(y, A, NX, NY, 50000, 0.0, l1, 0.0, ltsv, 1.0, nonneg=0, rel_tol=1e-14, trace=True)
    assert o["rc"] == 0
    assert (o["x"] < 0).any() and (o["x"] > 0).any()  # the signed prox is exercised
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-9)
    assert np.abs(o["x"].ravel() - xb).max() < 1e-4
    assert np.all(np.diff(o["cost"]) <= 1e-12)


def test_dft_signed_l1_box_reaches_optimum():
    """Signed L1 with a CLEAN box (soft_threshold_box: boxed pixels are 0)."""
    NX = NY = 8
    A, y = _dft_problem(62, NX, NY)
    box = np.ones(NX * NY, np.float32)
    box[::3] = 0.0
    l1, ltsv = 0.05, 0.2

    def smooth(x):
        r = y - A @ x
        t, dt = _tsv_fg(x, NX, NY)
        return 0.5 * r @ r + ltsv * t, -A.T @ r + ltsv * dt

    xb, fb = _solve_split(smooth, NX * NY, l1, fixed=(box == 0))
    o = oracle.imaging_core_dft(y, A, NX, NY, 50000, 0.0, l1, 0.0, ltsv, 1.0, nonneg=0, cl_box=box, rel_tol=1e-14)
    assert o["rc"] == 0
    assert np.all(o["x"].ravel()[box == 0] == 0)
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-9)


def _tv_smooth(x, NX, NY, eps):
    """Isotropic TV in the x[NX*j+i] view (c/mfista_TV_lib.c:30-46), smoothed
    by eps, and its gradient."""
    im = x.reshape(NY, NX)
    dh = im[:, :-1] - im[:, 1:]
    dv = im[:-1, :] - im[1:, :]
    inner = np.sqrt(dh[:-1, :] ** 2 + dv[:, :-1] ** 2 + eps)
    lr = np.sqrt(dh[-1, :] ** 2 + eps)
    lc = np.sqrt(dv[:, -1] ** 2 + eps)
    g = np.zeros_like(im)
    a, c = dh[:-1, :] / inner, dv[:, :-1] / inner
    g[:-1, :-1] += a + c
    g[:-1, 1:] -= a
    g[1:, :-1] -= c
    t = dh[-1, :] / lr
    g[-1, :-1] += t
    g[-1, 1:] -= t
    t = dv[:, -1] / lc
    g[:-1, -1] += t
    g[1:, -1] -= t
    return inner.sum() + lr.sum() + lc.sum(), g.ravel()


def test_dft_signed_l1_tv_close_to_smoothed_optimum():
    """L1+TV with nonneg_flag = 0: the TV prox is FGP_L1 (the signed soft
    threshold inside the dual iterations); the converged objective must match
    the finely smoothed optimum to 1e-5."""
    rng = np.random.default_rng(23)
    NX = NY = 6
    N, M = NX * NY, 50
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    xs = rng.standard_normal(N)
    y = A @ xs + 0.05 * rng.standard_normal(M)
    l1, ltv, eps = 0.05, 0.1, 1e-10

    def smooth(x):
        r = y - A @ x
        tv, gt = _tv_smooth(x, NX, NY, eps)
        return 0.5 * r @ r + ltv * tv, -A.T @ r + ltv * gt

    xb, fb = _solve_split(smooth, N, l1)
    o = oracle.imaging_core_dft(y, A, NX, NY, 20000, 0.0, l1, ltv, 0.0, 1.0, nonneg=0, rel_tol=1e-13, fgp_iter=300)
    assert o["rc"] == 0
    assert (o["x"] < 0).any()
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-5)


def test_dft_tv_box_close_to_smoothed_optimum():
    """L1+TV, nonneg, with a CLEAN box (FGP_nonneg_box, P_positive_box)."""
    rng = np.random.default_rng(24)
    NX = NY = 6
    N, M = NX * NY, 50
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    y = A @ np.maximum(rng.standard_normal(N), 0) + 0.05 * rng.standard_normal(M)
    box = np.ones(N, np.float32)
    box[:8] = 0.0
    l1, ltv, eps = 0.05, 0.1, 1e-10

    def fg(x):
        r = y - A @ x
        tv, gt = _tv_smooth(x, NX, NY, eps)
        return 0.5 * r @ r + l1 * x.sum() + ltv * tv, -A.T @ r + l1 + ltv * gt

    bounds = [(0, 0) if box[k] == 0 else (0, None) for k in range(N)]
    r = minimize(fg, np.zeros(N), jac=True, method="L-BFGS-B", bounds=bounds,
                 options={"maxiter": 50000, "maxfun": 100000, "ftol": 1e-16, "gtol": 1e-11})
    o = oracle.imaging_core_dft(y, A, NX, NY, 20000, 0.0, l1, ltv, 0.0, 1.0, nonneg=1, cl_box=box, rel_tol=1e-13,
                                fgp_iter=300)
    assert o["rc"] == 0
    assert np.all(o["x"].ravel()[box == 0] == 0)
    assert o["result"]["finalcost"] == pytest.approx(r.fun, rel=1e-5)
