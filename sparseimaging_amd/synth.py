"""Seeded synthetic skies and (u,v) coverages for the benchmark configurations.

Conventions follow the reference data files (c/mfista_imaging_nufft.c:141-170):
u, v in radians per pixel; image x[i, j] with i <-> u and j <-> v; model
visibility V(u, v) = sum_ij x_ij exp(+i (u (i - N/2) + v (j - N/2))).
Every generator takes an explicit seed (numpy PCG64).
"""
from __future__ import annotations

import numpy as np


def gaussian_blobs(N, rng, n, sig=(2.0, 8.0), amp=(0.2, 1.0), lo=None, hi=None, elliptical=False):
    lo = N / 4 if lo is None else lo
    hi = 3 * N / 4 if hi is None else hi
    i = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    x = np.zeros((N, N))
    for _ in range(n):
        ci, cj = rng.uniform(lo, hi, 2)
        a = rng.uniform(*amp)
        if elliptical:
            s1, s2 = rng.uniform(*sig, 2)
            th = rng.uniform(0, np.pi)
            di, dj = i - ci, j - cj
            p = di * np.cos(th) + dj * np.sin(th)
            q = -di * np.sin(th) + dj * np.cos(th)
            x += a * np.exp(-0.5 * (p * p / (s1 * s1) + q * q / (s2 * s2)))
        else:
            s = rng.uniform(*sig)
            x += a * np.exp(-0.5 * ((i - ci) ** 2 + (j - cj) ** 2) / (s * s))
    return x


def ring(N, radius, fwhm, ci=None, cj=None):
    ci = N / 2 if ci is None else ci
    cj = N / 2 if cj is None else cj
    i = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    r = np.hypot(i - ci, j - cj)
    s = fwhm / (2 * np.sqrt(2 * np.log(2)))
    return np.exp(-0.5 * ((r - radius) / s) ** 2)


def elliptical_arcs(M, rng, narcs=12, amin=0.05, amax=0.45, rot=0.0):
    """EHT-like tracks: `narcs` elliptical arcs with semi-major axis in
    [amin, amax] cycles/px; returns u, v in radians/px."""
    counts = np.full(narcs, M // narcs)
    counts[: M - counts.sum()] += 1
    us, vs = [], []
    for n in counts:
        a = rng.uniform(amin, amax)
        b = a * rng.uniform(0.3, 1.0)
        th = rng.uniform(0, np.pi) + rot
        ph0 = rng.uniform(0, 2 * np.pi)
        span = rng.uniform(np.pi / 3, np.pi)
        ph = ph0 + span * np.arange(n) / max(n, 1)
        uc = a * np.cos(ph) * np.cos(th) - b * np.sin(ph) * np.sin(th)
        vc = a * np.cos(ph) * np.sin(th) + b * np.sin(ph) * np.cos(th)
        us.append(uc)
        vs.append(vc)
    return 2 * np.pi * np.concatenate(us), 2 * np.pi * np.concatenate(vs)


def disc_1_over_r(M, rng, rmax=0.48):
    """Uniform radius in [0, rmax] cycles/px (density ~ 1/r), uniform angle."""
    r = rmax * rng.uniform(0, 1, M)
    t = rng.uniform(0, 2 * np.pi, M)
    return 2 * np.pi * r * np.cos(t), 2 * np.pi * r * np.sin(t)


def dft_vis(x, u, v, chunk=4096):
    """Exact FP64 V_k = sum_ij x_ij exp(+i(u_k (i-N/2) + v_k (j-N/2)))."""
    NX, NY = x.shape
    p = np.arange(NX) - NX // 2
    q = np.arange(NY) - NY // 2
    out = np.empty(u.size, dtype=np.complex128)
    for s in range(0, u.size, chunk):
        uu, vv = u[s:s + chunk], v[s:s + chunk]
        Ey = np.exp(1j * vv[:, None] * q[None, :])  # m x NY
        B = Ey @ x.T.astype(np.complex128)  # m x NX : sum_j x[i,j] e^{i v q_j}
        Ex = np.exp(1j * uu[:, None] * p[None, :])
        out[s:s + chunk] = np.sum(Ex * B, axis=1)
    return out


def make_nufft_problem(N=256, M=20000, seed=1002, kind="ring", noise_frac=0.01, narcs=12,
                       amax=0.45, coverage="arcs"):
    rng = np.random.default_rng(seed)
    if kind == "ring":
        x = ring(N, 40.0 * N / 256, 6.0 * N / 256) + gaussian_blobs(N, rng, 3)
    elif kind == "tv":
        x = gaussian_blobs(N, rng, 20) + ring(N, N * 0.15, 4.0) + ring(N, N * 0.3, 6.0, N * 0.45, N * 0.55)
    else:
        x = gaussian_blobs(N, rng, 5, elliptical=True)
    x = np.maximum(x, 0.0)
    x /= x.max()
    if coverage == "arcs":
        u, v = elliptical_arcs(M, rng, narcs=narcs, amax=amax)
    else:
        u, v = disc_1_over_r(M, rng, rmax=amax)
    vis = dft_vis(x, u, v)
    sig = noise_frac * np.abs(vis).max()
    vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    return {
        "u": u, "v": v, "vis_r": vis.real.copy(), "vis_i": vis.imag.copy(),
        "std": np.full(M, sig), "x_true": x, "N": N, "M": M, "seed": seed,
    }


def make_fft_problem(N=64, ncell=1500, seed=1001, noise_frac=0.01):
    """FFT-mode data: `ncell` distinct integer cells in the half-plane with
    |u|,|v| <= N/2-1, plus their conjugate partners (c/mfista_fft_lib.c:54-57)."""
    rng = np.random.default_rng(seed)
    x = gaussian_blobs(N, rng, 5, sig=(2.0, 8.0), elliptical=True, lo=N / 4, hi=3 * N / 4)
    x = np.maximum(x, 0.0)
    x /= x.max()
    h = N // 2 - 1
    cells = [(a, b) for a in range(-h, h + 1) for b in range(0, h + 1) if (b > 0 or a > 0)]
    pick = rng.choice(len(cells), size=ncell, replace=False)
    F = np.fft.fft2(x) / np.sqrt(N * N)
    us, vs, yr, yi = [], [], [], []
    Fsel = np.array([F[cells[k][0] % N, cells[k][1] % N] for k in pick])
    sig = noise_frac * np.abs(Fsel).max()
    noise = sig * (rng.standard_normal(ncell) + 1j * rng.standard_normal(ncell))
    for t, k in enumerate(pick):
        a, b = cells[k]
        val = Fsel[t] + noise[t]
        us += [a % N, (-a) % N]
        vs += [b % N, (-b) % N]
        yr += [val.real, val.real]
        yi += [val.imag, -val.imag]
    M = len(us)
    return {
        "u_idx": np.array(us, dtype=np.int32), "v_idx": np.array(vs, dtype=np.int32),
        "y_r": np.array(yr), "y_i": np.array(yi), "noise": np.full(M, sig), "x_true": x,
        "N": N, "M": M, "seed": seed,
    }


def make_movie_frames(nframes=32, N=128, seed=1004, mmin=5000, mmax=40000, drift=0.5, noise_frac=0.01):
    """Movie of a drifting ring: frame f is a ring of radius 0.16 N (FWHM
    0.05 N) whose centre moves `drift` px per frame along the diagonal, plus
    2 fixed Gaussians; its (u,v) coverage is 12 elliptical arcs whose rotation
    advances with the frame index, with M_f drawn uniformly from
    {mmin .. mmax}.  Returns a list of make_nufft_problem-style dicts."""
    rng = np.random.default_rng(seed)
    blobs = gaussian_blobs(N, rng, 2)
    frames = []
    for f in range(nframes):
        M = int(rng.integers(mmin, mmax + 1))
        c = N / 2 - drift * nframes / 2 + drift * f
        x = np.maximum(ring(N, 0.16 * N, 0.05 * N, c, c) + blobs, 0.0)
        x /= x.max()
        u, v = elliptical_arcs(M, rng, rot=np.pi * f / nframes)
        vis = dft_vis(x, u, v)
        sig = noise_frac * np.abs(vis).max()
        vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
        frames.append({"u": u, "v": v, "vis_r": vis.real.copy(), "vis_i": vis.imag.copy(),
                       "std": np.full(M, sig), "x_true": x, "N": N, "M": M, "seed": seed, "frame": f})
    return frames
