"""Pins the oracle's MFISTA loop against an independent optimiser (CPU only).

With nonneg=1 the L1 term is linear (sum x), so L1+TSV is a smooth,
bound-constrained convex problem that scipy's L-BFGS-B solves independently
of the oracle; the converged MFISTA objective must reach the same optimum.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle


def _lbfgsb(fun_grad, n):
    r = minimize(fun_grad, np.zeros(n), jac=True, method="L-BFGS-B", bounds=[(0, None)] * n,
                 options={"maxiter": 50000, "maxfun": 100000, "ftol": 1e-16, "gtol": 1e-11})
    return r.x, r.fun


def _tsv_fg(x, NX, NY):
    return oracle.tsv_flat(x, NX, NY), oracle.d_tsv_flat(x, NX, NY)


def test_dft_l1_tsv_reaches_optimum():
    rng = np.random.default_rng(21)
    NX = NY = 8
    N, M = NX * NY, 90
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    xt = np.maximum(rng.standard_normal(N), 0)
    y = A @ xt + 0.05 * rng.standard_normal(M)
    l1, ltsv = 0.05, 0.2

    def fg(x):
        r = y - A @ x
        t, dt = _tsv_fg(x, NX, NY)
        return 0.5 * r @ r + l1 * x.sum() + ltsv * t, -A.T @ r + l1 + ltsv * dt

    xb, fb = _lbfgsb(fg, N)
    o = oracle.imaging_core_dft(y, A, NX, NY, 50000, 0.0, l1, 0.0, ltsv, 1.0, nonneg=1, rel_tol=1e-14,
                                trace=True)
    assert o["rc"] == 0
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-9)
    assert np.abs(o["x"].ravel() - xb).max() < 1e-4
    assert np.all(np.diff(o["cost"]) <= 1e-12)  # monotone FISTA
    # RESULT fields (c/mfista_dft_lib.c:440-536)
    res = o["result"]
    r = y - A @ o["x"].ravel()
    assert res["sq_error"] == pytest.approx(r @ r, rel=1e-12)
    assert np.allclose(o["residual"], r)
    assert res["M"] == M and res["N"] == N


@pytest.mark.parametrize("preset", [oracle.PRESET_C24, oracle.PRESET_CXX12])
def test_nufft_l1_tsv_reaches_optimum(preset):
    rng = np.random.default_rng(31)
    N, M = 32, 300
    u = rng.uniform(-1.0, 1.0, M)
    v = rng.uniform(-1.0, 1.0, M)
    xt = np.zeros((N, N))
    xt[12:20, 10:22] = 1.0
    vis = oracle.nufft_forward(u, v, xt, preset) + 0.05 * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    sd = np.full(M, 0.05)
    w2 = 1 / sd ** 2
    l1, ltsv = 1.0, 5.0

    def fg(x):
        Ax = oracle.nufft_forward(u, v, x.reshape(N, N), preset)
        r = vis - Ax
        g = -oracle.nufft_adjoint(u, v, w2 * r, N, N, preset).ravel()
        t, dt = _tsv_fg(x, N, N)
        return 0.5 * np.sum(w2 * np.abs(r) ** 2) + l1 * x.sum() + ltsv * t, g + l1 + ltsv * dt

    xb, fb = _lbfgsb(fg, N * N)
    o = oracle.imaging_core_nufft(u, v, vis.real, vis.imag, sd, N, N, 20000, 0.0, l1, 0.0, ltsv, 1e4,
                                  nonneg=1, preset=preset, rel_tol=1e-13, trace=True)
    assert o["rc"] == 0
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-8)
    assert np.all(np.diff(o["cost"]) <= 1e-9 * o["cost"][0])
    assert o["result"]["M"] == (M // 2 if preset == oracle.PRESET_C24 else M)


def test_fft_mode_l1_tsv_reaches_optimum():
    rng = np.random.default_rng(41)
    N = 16
    cells = {(int(a), int(b)) for a, b in zip(rng.integers(-7, 8, 80), rng.integers(0, 8, 80)) if b > 0 or a > 0}
    u, v, yr, yi = [], [], [], []
    for a, b in sorted(cells):
        val = rng.standard_normal() + 1j * rng.standard_normal()
        u += [a % N, (-a) % N]
        v += [b % N, (-b) % N]
        yr += [val.real, val.real]
        yi += [val.imag, -val.imag]
    u, v = np.array(u, np.int32), np.array(v, np.int32)
    yr, yi = np.array(yr), np.array(yi)
    sd = np.full(u.size, 0.3)
    l1, ltsv = 0.2, 1.0

    def fg(x):
        F, g = oracle.fft_F_grad(u, v, yr, yi, sd, x.reshape(N, N))
        t, dt = _tsv_fg(x, N, N)
        return F + l1 * x.sum() + ltsv * t, -g.ravel() + l1 + ltsv * dt

    xb, fb = _lbfgsb(fg, N * N)
    o = oracle.imaging_core_fft(u, v, yr, yi, sd, N, N, 50000, 0.0, l1, 0.0, ltsv, 10.0, nonneg=1,
                                rel_tol=1e-14)
    assert o["rc"] == 0
    assert o["result"]["finalcost"] == pytest.approx(fb, rel=1e-9)


def test_dft_l1_tv_close_to_smoothed_optimum():
    """TV through FGP: the MFISTA objective must not exceed the optimum of a
    finely smoothed TV (eps=1e-10) by more than 1e-5 relative."""
    rng = np.random.default_rng(22)
    NX = NY = 6
    N, M = NX * NY, 50
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    y = A @ np.maximum(rng.standard_normal(N), 0) + 0.05 * rng.standard_normal(M)
    l1, ltv = 0.05, 0.1
    eps = 1e-10

    def fg(x):
        r = y - A @ x
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
        tv = inner.sum() + lr.sum() + lc.sum()
        return 0.5 * r @ r + l1 * x.sum() + ltv * tv, -A.T @ r + l1 + ltv * g.ravel()

    xb, fb = _lbfgsb(fg, N)
    o = oracle.imaging_core_dft(y, A, NX, NY, 20000, 0.0, l1, ltv, 0.0, 1.0, nonneg=1, rel_tol=1e-13,
                                fgp_iter=300)
    assert o["rc"] == 0
    assert o["result"]["finalcost"] <= fb * (1 + 1e-5)
    assert o["result"]["finalcost"] >= fb * (1 - 1e-5)


def test_error_paths():
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, 20)
    v = rng.uniform(-1, 1, 20)
    ones = np.ones(20)
    o = oracle.imaging_core_nufft(u, v, ones, ones, ones, 32, 32, 10, 1e-5, 1.0, 1.0, 1.0, 1.0)
    assert o["rc"] == -1  # both regularisers: RESULT untouched
    assert o["result"]["ITER"] == 0 and o["result"]["finalcost"] == 0.0
    o = oracle.imaging_core_nufft(u, v, ones, ones, ones, 32, 32, 10, 1e-5, 1.0, 0.0, 1.0, 1.0, nonneg=2)
    assert o["rc"] == 0 and o["result"]["ITER"] == 0  # c/mfista_nufft_lib.c:634-637
    o = oracle.imaging_core_nufft(u, v, ones, ones, ones, 32, 32, 10, 1e-5, 1.0, 1.0, 0.0, 1.0,
                                  preset=oracle.PRESET_CXX12)
    assert o["rc"] == -2  # TV is C-library only


def test_absolute_stopping_rule_uses_scaled_eps():
    """eps is multiplied by mean |V|^2 (c/mfista_nufft_lib.c:923-926): scaling
    the data by s (and sigma by s) leaves chi^2 unchanged but multiplies
    the threshold by s^2, so the solve stops no later."""
    rng = np.random.default_rng(2)
    N, M = 32, 200
    u = rng.uniform(-1, 1, M)
    v = rng.uniform(-1, 1, M)
    vis = oracle.nufft_forward(u, v, np.maximum(rng.standard_normal((N, N)), 0), 0)
    sd = np.full(M, 1.0)
    a = oracle.imaging_core_nufft(u, v, vis.real, vis.imag, sd, N, N, 3000, 1e-7, 1.0, 0.0, 1.0, 1e3)
    b = oracle.imaging_core_nufft(u, v, 10 * vis.real, 10 * vis.imag, 10 * sd, N, N, 3000, 1e-7, 10.0,
                                  0.0, 100.0, 1e1)
    assert b["result"]["ITER"] <= a["result"]["ITER"]
