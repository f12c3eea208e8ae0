"""The coupled movie oracle (oracle/movie_ref.py, a restatement of
matlab/MFISTA_movie_eht.m) pinned on the CPU:

- d_X and dt_X against closed forms and finite differences;
- the frame times of a two-column time_idx (mean time per frame, rescaled);
- lambda_t = 0: the frames decouple, so each converged frame must be the
  single-frame optimum (the single-frame oracle MFISTA, itself pinned by
  L-BFGS-B in test_oracle_mfista.py);
- lambda_t > 0: the converged objective must match an independent L-BFGS-B
  optimum of the full coupled objective (nonneg makes L1 linear, so the
  problem is smooth and bound-constrained)."""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from oracle import movie_ref as mr


def _frames(nt, N, M, seed):
    rng = np.random.default_rng(seed)
    frames, truth = [], []
    yy, xx = np.mgrid[0:N, 0:N]
    for t in range(nt):
        img = np.exp(-((xx - N / 2 - 0.8 * t) ** 2 + (yy - N / 2) ** 2) / (2 * (N / 8) ** 2))
        r = 0.28 * np.sqrt(rng.uniform(0, 1, M))  # cycles/px; 24-tap footprint inside the 2N grid at N=32
        th = rng.uniform(0, 2 * np.pi, M)
        u, v = 2 * np.pi * r * np.cos(th), 2 * np.pi * r * np.sin(th)
        vis = oracle.nufft_forward(u, v, img)
        sig = 0.02 * np.abs(vis).max()
        vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
        frames.append({"u": u, "v": v, "vis": vis, "std": np.full(M, sig)})
        truth.append(img)
    return frames, np.array(truth)


def test_d_X_closed_form_and_gradient():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 3, 5))
    t = np.array([0.0, 1.0, 3.0, 3.5])
    w = mr.time_weights(t)
    assert np.allclose(w, [1.0, 0.5, 2.0])
    ref = sum(w[k] * np.sum((x[k] - x[k + 1]) ** 2) for k in range(3))
    assert abs(mr.d_X(x, w) - ref) <= 1e-12 * ref
    g = mr.dt_X(x, w)
    h = 1e-6
    for idx in [(0, 0, 0), (1, 2, 4), (3, 1, 2), (2, 0, 3)]:
        e = np.zeros_like(x)
        e[idx] = h
        fd = (mr.d_X(x + e, w) - mr.d_X(x - e, w)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(g[idx]))
    # a constant movie has no temporal cost and no temporal gradient
    c = np.repeat(x[:1], 4, axis=0)
    assert mr.d_X(c, w) == 0.0 and np.all(mr.dt_X(c, w) == 0.0)


def test_frame_times_from_time_column():
    ti = np.array([[1, 0.0], [1, 2.0], [2, 5.0], [3, 9.0], [3, 11.0]])
    t = mr.frame_times(ti, 3)
    mean = np.array([1.0, 5.0, 10.0])
    assert np.allclose(t, 3 * mean / 9.0)
    assert np.allclose(mr.frame_times(ti[:, 0], 3), [1, 2, 3])


def test_lambda_t_zero_decouples_frames():
    N = 32
    frames, _ = _frames(2, N, 150, 11)
    w = mr.time_weights([1.0, 2.0])
    lam, lam2 = 0.05, 0.5
    o = mr.movie_mfista(frames, N, N, np.zeros((2, N, N)), lam, lam2, 0.0, w, True, 1e3, 20000, 1e-12)
    for k, fr in enumerate(frames):
        s = oracle.imaging_core_nufft(fr["u"], fr["v"], fr["vis"].real, fr["vis"].imag, fr["std"], N, N, 20000, 0.0,
                                      lam, 0.0, lam2, 1e3, nonneg=1, rel_tol=1e-12)
        fk = mr.objective([fr], o["x"][k:k + 1], lam, lam2, 0.0, w[:0])
        fs = mr.objective([fr], s["x"][None], lam, lam2, 0.0, w[:0])
        e = np.linalg.norm(o["x"][k] - s["x"]) / np.linalg.norm(s["x"])
        assert abs(fk - fs) <= 1e-7 * fs, (k, fk, fs)
        assert e <= 1e-3, (k, e)


def _lbfgs(frames, N, lam, lam2, lam_t, w):
    nt = len(frames)

    def fg(z):
        x = z.reshape(nt, N, N)
        f = 0.0
        g = np.zeros_like(x)
        for k, fr in enumerate(frames):
            r = (fr["vis"] - oracle.nufft_forward(fr["u"], fr["v"], x[k])) / fr["std"]
            f += 0.5 * float(np.real(np.vdot(r, r))) + lam2 * mr.tsv(x[k]) + lam * x[k].sum()
            g[k] = -oracle.nufft_adjoint(fr["u"], fr["v"], r / fr["std"], N, N) + lam2 * mr.d_tsv(x[k]) + lam
        f += lam_t * mr.d_X(x, w)
        g += lam_t * mr.dt_X(x, w)
        return f, g.ravel()

    res = minimize(fg, np.zeros(nt * N * N), jac=True, method="L-BFGS-B", bounds=[(0, None)] * (nt * N * N),
                   options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-11, "maxcor": 30})
    return res.x.reshape(nt, N, N), res.fun


@pytest.mark.parametrize("lam_t", [2.0, 50.0])
def test_coupled_optimum_matches_lbfgs(lam_t):
    N = 32
    frames, _ = _frames(3, N, 150, 12)
    w = mr.time_weights([1.0, 2.0, 4.0])
    lam, lam2 = 0.05, 0.5
    o = mr.movie_mfista(frames, N, N, np.zeros((3, N, N)), lam, lam2, lam_t, w, True, 1e3, 20000, 1e-12)
    xl, fl = _lbfgs(frames, N, lam, lam2, lam_t, w)
    fo = mr.objective(frames, o["x"], lam, lam2, lam_t, w)
    e = np.linalg.norm(o["x"] - xl) / np.linalg.norm(xl)
    print(f"lambda_t={lam_t}: MFISTA iters {o['iters']} F {fo:.12e} L-BFGS-B {fl:.12e} image relL2 {e:.2e}")
    assert abs(fo - fl) <= 1e-8 * abs(fl)
    assert e <= 1e-3
    # monotone objective history, and cost[0] is the objective of the start
    assert np.all(np.diff(o["cost"]) <= 1e-9 * abs(o["cost"][0]))
    assert abs(o["cost"][0] - mr.objective(frames, np.zeros((3, N, N)), lam, lam2, lam_t, w)) <= 1e-9 * o["cost"][0]
    # a larger lambda_t pulls the frames together
    if lam_t > 10:
        o2 = mr.movie_mfista(frames, N, N, np.zeros((3, N, N)), lam, lam2, 0.1, w, True, 1e3, 20000, 1e-12)
        assert mr.d_X(o["x"], w) < 0.5 * mr.d_X(o2["x"], w)
