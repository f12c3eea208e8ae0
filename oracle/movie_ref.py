"""FP64 CPU restatement of the coupled movie solver -- TEST INFRASTRUCTURE ONLY.

Follows matlab/MFISTA_movie_eht.m:1-239 (with matlab/lib/d_X.m, dt_X.m,
d_TSV.m, TSV.m, softth.m, softth_nonneg.m, costF_nufft_TSV.m): MFISTA over
the stack of N_t frames with one step-size constant c and one momentum
sequence mu, minimising

    sum_t |W_t (V_t - A_t x_t)|^2 / 2 + lambda sum_t |x_t|_1
        + lambda2 sum_t TSV(x_t) + lambda_t d_X(x),
    d_X(x) = sum_{t < N_t} w_t |x_t - x_{t+1}|^2,   w_t = 1 / (t_{t+1} - t_t).

The reference computes A_t with FINUFFT at 1e-12 (type 2, isign +1, modes
-N/2 .. N/2-1; MFISTA_movie_eht.m:95-96); here A_t is the oracle's FP64
24-tap NUFFT (mfista_oracle.c, C-24 preset, 2e-11 of the exact DFT), the same
operator definition.  Image vectors are (Nx, Ny) arrays with the first index
along u, as the reference's column-major reshape(x(:, i), Nx, Ny).

Pinned (tests/test_oracle_movie.py) by: the d_X / dt_X closed forms and finite
differences; lambda_t = 0 reducing to independent frames (each converged
frame matches the single-frame oracle MFISTA optimum); and the converged
objective against an independent L-BFGS-B optimum of the full coupled
objective.
"""
from __future__ import annotations

import numpy as np

from . import PRESET_C24, nufft_adjoint, nufft_forward

MINITER, TD, ETA = 100, 50, 1.1  # MFISTA_movie_eht.m:29-32


def time_weights(t_list):
    """w_t = 1 / (t_{t+1} - t_t) (matlab/lib/d_X.m:6)."""
    t = np.asarray(t_list, dtype=np.float64)
    return 1.0 / (t[1:] - t[:-1])


def frame_times(time_idx, nframes):
    """t_list of MFISTA_movie_eht.m:53-69: the mean observing time of each
    frame rescaled to span N_t when time_idx has a second column, else 1..N_t."""
    ti = np.asarray(time_idx)
    if ti.ndim == 2 and ti.shape[1] == 2:
        fr = ti[:, 0].astype(int) - 1
        s = np.bincount(fr, weights=ti[:, 1].astype(np.float64), minlength=nframes)
        n = np.bincount(fr, minlength=nframes)
        t = s / n
        return nframes * t / (t.max() - t.min())
    return np.arange(1, nframes + 1, dtype=np.float64)


def d_X(x, w):
    """matlab/lib/d_X.m: sum_t w_t |x_t - x_{t+1}|^2 for x of shape (N_t, Nx, Ny)."""
    d = x[:-1] - x[1:]
    return float(np.sum(w[:, None, None] * d * d))


def dt_X(x, w):
    """matlab/lib/dt_X.m: the gradient of d_X."""
    g = np.zeros_like(x)
    tmp = 2.0 * (x[:-1] - x[1:]) * w[:, None, None]
    g[:-1] += tmp
    g[1:] -= tmp
    return g


def tsv(x):
    """matlab/lib/TSV.m on one (Nx, Ny) frame."""
    return float(np.sum((x[:-1, :] - x[1:, :]) ** 2) + np.sum((x[:, :-1] - x[:, 1:]) ** 2))


def d_tsv(x):
    """matlab/lib/d_TSV.m on one (Nx, Ny) frame."""
    g = np.zeros_like(x)
    a = 2.0 * (x[:-1, :] - x[1:, :])
    b = 2.0 * (x[:, :-1] - x[:, 1:])
    g[:-1, :] += a
    g[1:, :] -= a
    g[:, :-1] += b
    g[:, 1:] -= b
    return g


def _prox(v, eta, nonneg, box):
    """softth / softth_nonneg (matlab/lib/softth*.m), times the CLEAN box."""
    if nonneg:
        out = np.where(v - eta > 0, v - eta, 0.0)
    else:
        out = np.where(np.abs(v) - eta > 0, v - np.sign(v) * eta, 0.0)
    return out if box is None else out * box


def objective(frames, x, lam, lam2, lam_t, w, preset=PRESET_C24):
    """The coupled objective; frames is a list of dicts u, v, vis (complex), std."""
    f = 0.0
    for k, fr in enumerate(frames):
        r = (fr["vis"] - nufft_forward(fr["u"], fr["v"], x[k], preset)) / fr["std"]
        f += 0.5 * float(np.real(np.vdot(r, r))) + lam2 * tsv(x[k]) + lam * float(np.abs(x[k]).sum())
    return f + lam_t * d_X(x, w)


def movie_mfista(frames, Nx, Ny, xinit, lam, lam2, lam_t, w, nonneg, cinit, maxiter, eps, box=None,
                 preset=PRESET_C24):
    """MFISTA_movie_eht.m:44-169.  xinit: (N_t, Nx, Ny) (the reference repeats
    one image, :49); box: (Nx, Ny) real mask or None.  Returns dict x, cost
    (per iteration), iters, c."""
    nt = len(frames)
    wv = [fr["vis"] / fr["std"] for fr in frames]  # data(i).wvis, :80
    sinv = [1.0 / fr["std"] for fr in frames]
    x = np.array(xinit, dtype=np.float64, copy=True)
    if box is not None:
        x = x * box[None]
    z = x.copy()
    c = cinit
    mu = 1.0

    def fwd(k, img):
        return nufft_forward(frames[k]["u"], frames[k]["v"], img, preset)

    tmpc = 0.0  # :91-98
    for k in range(nt):
        yax = wv[k] - fwd(k, x[k]) * sinv[k]
        tmpc += lam * float(np.abs(x[k]).sum()) + 0.5 * float(np.real(np.vdot(yax, yax))) + lam2 * tsv(x[k])
    # The reference's initial cost has no temporal term (:91-98); it starts
    # every frame from the same image (:49), where d_X = 0, so adding the term
    # changes nothing there and keeps cost[0] the objective for any start.
    tmpc += lam_t * d_X(x, w)
    cost = []
    it = 0
    for it in range(maxiter):  # :102
        cost.append(tmpc)
        Qcore = 0.0
        grad = np.zeros_like(x)
        for k in range(nt):  # :112-131
            yaz = wv[k] - fwd(k, z[k]) * sinv[k]
            Qcore += 0.5 * float(np.real(np.vdot(yaz, yaz))) + lam2 * tsv(z[k])
            grad[k] = nufft_adjoint(frames[k]["u"], frames[k]["v"], yaz * sinv[k], Nx, Ny, preset) - lam2 * d_tsv(z[k])
        Qcore += lam_t * d_X(z, w)  # :133
        grad -= lam_t * dt_X(z, w)  # :135-137
        for _ in range(maxiter):  # :139-169
            xtmp = _prox(grad / c + z, lam / c, nonneg, box)
            F = 0.0
            Q = Qcore
            for k in range(nt):
                yax = wv[k] - fwd(k, xtmp[k]) * sinv[k]
                F += 0.5 * float(np.real(np.vdot(yax, yax))) + lam2 * tsv(xtmp[k])
                dk = xtmp[k] - z[k]
                Q += -float(np.sum(dk * grad[k])) + float(np.sum(dk * dk)) * c / 2
            F += lam_t * d_X(xtmp, w)
            if F <= Q:
                break
            c *= ETA
            if c > 1.0e200:
                break
        c /= ETA  # :171
        munew = (1 + np.sqrt(1 + 4 * mu * mu)) / 2
        F += lam * float(np.abs(xtmp).sum())
        if F < cost[it]:  # :177-184
            tmpc = F
            z = (1 + (mu - 1) / munew) * xtmp + ((1 - mu) / munew) * x
            x = xtmp
        else:
            z = (mu / munew) * xtmp + (1 - (mu / munew)) * x
        if it + 1 > MINITER and (cost[it - TD] - cost[it]) < eps:  # :186 (1-based t > MINITER)
            break
        mu = munew
    return {"x": x, "cost": np.array(cost), "iters": it + 1, "c": c}
