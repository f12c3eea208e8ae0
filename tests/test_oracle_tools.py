"""Pins the oracle's elementwise tools, TSV/TV and the FGP TV prox (CPU only)."""
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(__file__), "golden")


def test_tsv_tv_dtsv_vs_golden_nonsquare():
    d = np.load(os.path.join(G, "reg_5x7.npz"))
    NX, NY = int(d["NX"]), int(d["NY"])
    assert abs(oracle.tsv_flat(d["x"], NX, NY) - float(d["tsv"])) < 1e-12
    assert abs(oracle.tv_flat(d["x"], NX, NY) - float(d["tv"])) < 1e-12
    assert np.abs(oracle.d_tsv_flat(d["x"], NX, NY) - d["dtsv"]).max() < 1e-12


def test_dtsv_finite_differences():
    rng = np.random.default_rng(3)
    NX, NY = 6, 9
    x = rng.standard_normal(NX * NY)
    g = oracle.d_tsv_flat(x, NX, NY)
    for f in range(NX * NY):
        e = np.zeros_like(x)
        e[f] = 1e-6
        fd = (oracle.tsv_flat(x + e, NX, NY) - oracle.tsv_flat(x - e, NX, NY)) / 2e-6
        assert abs(fd - g[f]) < 1e-6


def test_dtsv_hessian_diagonal_4_6_8():
    """TSV is quadratic: dTSV(e_f)[f] = 2 x (number of neighbours) = 4 at the
    corners, 6 on the edges, 8 inside (matlab/lib/d2_TSV.m)."""
    NX, NY = 5, 4
    for f in range(NX * NY):
        e = np.zeros(NX * NY)
        e[f] = 1.0
        i, j = f % NX, f // NX
        nb = (i > 0) + (i < NX - 1) + (j > 0) + (j < NY - 1)
        assert oracle.d_tsv_flat(e, NX, NY)[f] == pytest.approx(2 * nb)


def test_soft_thresholds_closed_form():
    v = np.array([-3.0, -1.0, -0.5, 0.0, 0.5, 1.0, 3.0])
    eta = 1.0
    assert np.array_equal(oracle.soft_threshold(v, eta), [-2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 2.0])
    assert np.array_equal(oracle.soft_threshold(v, eta, nonneg=True), [0, 0, 0, 0, 0, 0.0, 2.0])
    box = np.array([1, 1, 1, 1, 1, 1, 0], dtype=np.float32)
    assert np.array_equal(oracle.soft_threshold(v, eta, box=box), [-2.0, 0, 0, 0, 0, 0, 0])
    assert np.array_equal(oracle.soft_threshold(v, eta, nonneg=True, box=box), [0, 0, 0, 0, 0, 0, 0])


def test_calc_Q_part_closed_form():
    rng = np.random.default_rng(5)
    x1, x2, g = rng.standard_normal((3, 50))
    c = 3.7
    want = -np.dot(x1 - x2, g) + c * np.dot(x1 - x2, x1 - x2) / 2
    assert oracle.calc_Q_part(x1, x2, c, g) == pytest.approx(want, rel=1e-13)


def test_Lx_Lpq_are_adjoint():
    rng = np.random.default_rng(6)
    NX, NY = 7, 5
    x = rng.standard_normal(NX * NY)
    p = rng.standard_normal((NX - 1) * NY)
    q = rng.standard_normal(NX * (NY - 1))
    lp, lq = oracle.Lx_flat(x, NX, NY)
    assert np.dot(lp, p) + np.dot(lq, q) == pytest.approx(np.dot(x, oracle.Lpq_flat(p, q, NX, NY)), rel=1e-12)


def test_P_pqiso_closed_form():
    NX, NY = 3, 3
    p = np.array([3.0, 0.1, 0.2, 0.3, 5.0, -2.0])  # (NX-1)*NY
    q = np.array([4.0, 0.2, 0.5, 0.1, 0.1, 0.7])   # NX*(NY-1)
    r, s = oracle.P_pqiso_flat(p, q, NX, NY)
    # pixel (0,0): (3,4) -> norm 5 -> (0.6, 0.8)
    assert r[0] == pytest.approx(0.6) and s[0] == pytest.approx(0.8)
    # last row of p (j = NY-1): 1-D clamp |p| <= 1
    assert r[(NX - 1) * (NY - 1) + 0] == pytest.approx(1.0)
    assert r[(NX - 1) * (NY - 1) + 1] == pytest.approx(-1.0)
    # last column of q (i = NX-1): 0.5 and 0.7 stay
    assert s[NX * 0 + NX - 1] == pytest.approx(0.5) and s[NX * 1 + NX - 1] == pytest.approx(0.7)


def test_fgp_small_lambda_is_projection():
    rng = np.random.default_rng(7)
    NX = NY = 8
    b = rng.standard_normal(NX * NY)
    x = oracle.fgp(b, NX, NY, 0.0, 1e-9, 20, nonneg=True)
    assert np.abs(x - np.maximum(b, 0)).max() < 1e-6
    x = oracle.fgp(b, NX, NY, 0.3, 1e-9, 20, nonneg=False)
    assert np.abs(x - oracle.soft_threshold(b, 0.3)).max() < 1e-6


def test_fgp_constant_image_is_fixed_point():
    NX = NY = 6
    b = np.full(NX * NY, 2.5)
    assert np.abs(oracle.fgp(b, NX, NY, 0.0, 0.7, 30, nonneg=True) - 2.5).max() < 1e-12


def _tv_prox_bruteforce(b, NX, NY, lam):
    """min_x 1/2|x-b|^2 + lam*TV(x), x >= 0, with TV smoothed by eps=1e-9 (L-BFGS-B)."""
    from scipy.optimize import minimize

    eps = 1e-9

    def f(x):
        im = x.reshape(NY, NX)
        dh = im[:, :-1] - im[:, 1:]
        dv = im[:-1, :] - im[1:, :]
        inner = np.sqrt(dh[:-1, :] ** 2 + dv[:, :-1] ** 2 + eps)
        lr = np.sqrt(dh[-1, :] ** 2 + eps)
        lc = np.sqrt(dv[:, -1] ** 2 + eps)
        val = 0.5 * np.sum((x - b) ** 2) + lam * (inner.sum() + lr.sum() + lc.sum())
        g = np.zeros_like(im)
        a = dh[:-1, :] / inner
        c = dv[:, :-1] / inner
        g[:-1, :-1] += a
        g[:-1, 1:] -= a
        g[:-1, :-1] += c
        g[1:, :-1] -= c
        t = dh[-1, :] / lr
        g[-1, :-1] += t
        g[-1, 1:] -= t
        t = dv[:, -1] / lc
        g[:-1, -1] += t
        g[1:, -1] -= t
        return val, (x - b) + lam * g.ravel()

    r = minimize(f, np.maximum(b, 0), jac=True, method="L-BFGS-B", bounds=[(0, None)] * b.size,
                 options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-12})
    return r.x, r.fun


def test_fgp_nonneg_matches_bruteforce_4x4():
    rng = np.random.default_rng(8)
    NX = NY = 4
    b = rng.standard_normal(NX * NY) + 0.5
    lam = 0.3
    xb, fb = _tv_prox_bruteforce(b, NX, NY, lam)
    x = oracle.fgp(b, NX, NY, 0.0, lam, 3000, nonneg=True)
    obj = 0.5 * np.sum((x - b) ** 2) + lam * oracle.tv_flat(x, NX, NY)
    assert obj <= fb + 1e-6
    assert np.abs(x - xb).max() < 2e-3


def test_fgp_objective_decreases_with_iterations():
    rng = np.random.default_rng(10)
    NX = NY = 12
    b = rng.standard_normal(NX * NY)
    lam = 0.5

    def obj(x):
        return 0.5 * np.sum((x - b) ** 2) + lam * oracle.tv_flat(x, NX, NY)

    o5 = obj(oracle.fgp(b, NX, NY, 0.0, lam, 5, nonneg=True))
    o200 = obj(oracle.fgp(b, NX, NY, 0.0, lam, 200, nonneg=True))
    assert o200 < o5
