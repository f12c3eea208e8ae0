"""Coupled movie solver on the GPU (mfista_movie_nufft, one persistent launch
for all frames) against the movie oracle (oracle/movie_ref.py, pinned on the
CPU by L-BFGS-B) on the same inputs, both run to convergence.

Bar: image relative L2 <= 1e-3 and objective <= 1e-6 relative (FP32 adjoint,
FP64 iterates and objective), as for the single-frame NUFFT solve."""
import numpy as np
import pytest

import oracle
from oracle import movie_ref as mr
import sparseimaging_amd as sa

pytestmark = pytest.mark.gpu


def _movie(nt, N, M, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:N, 0:N]
    u, v, vis, sd, ti, tv = [], [], [], [], [], []
    for t in range(nt):
        img = np.exp(-((xx - N / 2 - 0.8 * t) ** 2 + (yy - N / 2) ** 2) / (2 * (N / 8) ** 2))
        m = M + 20 * t
        r = 0.28 * np.sqrt(rng.uniform(0, 1, m))
        th = rng.uniform(0, 2 * np.pi, m)
        uu, vv = 2 * np.pi * r * np.cos(th), 2 * np.pi * r * np.sin(th)
        y = oracle.nufft_forward(uu, vv, img)
        sig = 0.02 * np.abs(y).max()
        y = y + sig * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
        u.append(uu), v.append(vv), vis.append(y), sd.append(np.full(m, sig))
        ti.append(np.full(m, t + 1)), tv.append(1.3 * t + rng.uniform(-0.2, 0.2, m) + 0.1 * t * t)
    # interleave the frames' visibilities as an observation would list them
    order = rng.permutation(sum(len(a) for a in u))
    cat = lambda a: np.concatenate(a)[order]
    return {"u": cat(u), "v": cat(v), "vis": cat(vis), "std": cat(sd), "frame": cat(ti).astype(np.int32),
            "time": cat(tv)}


def _frames_of(D, nt):
    return [{"u": D["u"][D["frame"] == t + 1], "v": D["v"][D["frame"] == t + 1],
             "vis": D["vis"][D["frame"] == t + 1], "std": D["std"][D["frame"] == t + 1]} for t in range(nt)]


@pytest.mark.parametrize("lam_t,timed", [(2.0, False), (50.0, True)])
def test_movie_matches_oracle(lam_t, timed):
    nt, N = 3, 32
    D = _movie(nt, N, 150, 21)
    lam, lam2 = 0.05, 0.5
    ti = np.stack([D["frame"], D["time"]], axis=1) if timed else D["frame"]
    print(f"case movie {nt} frames {N}x{N}, lambda_t={lam_t}, time column={timed}", flush=True)
    x, cost, terms, outvis, info = sa.mfista_movie_nufft(D["vis"], D["u"], D["v"], D["std"], ti, None, N, N, None,
                                                         lam, lam2, lam_t, nonneg_flag=1, cinit=1e3,
                                                         maxiter=30000, eps=1e-12)
    w = mr.time_weights(mr.frame_times(ti, nt))
    frames = _frames_of(D, nt)
    o = mr.movie_mfista(frames, N, N, np.zeros((nt, N, N)), lam, lam2, lam_t, w, True, 1e3, 30000, 1e-12)
    fg = mr.objective(frames, x, lam, lam2, lam_t, w)
    fo = mr.objective(frames, o["x"], lam, lam2, lam_t, w)
    e = np.linalg.norm(x - o["x"]) / np.linalg.norm(o["x"])
    print(f"  GPU: {info['iterations']} iterations, {info['total_trials']} trials, {info['trial_rounds']} rounds, "
          f"{info['kernel_ms']:.1f} ms; oracle {o['iters']} iterations; objective rel {abs(fg - fo) / fo:.2e}, "
          f"image relL2 {e:.2e}", flush=True)
    assert abs(fg - fo) <= 1e-6 * fo
    assert e <= 1e-3
    # outputs: cost history, terms and model visibilities describe the returned images
    assert len(cost) == info["iterations"] and np.all(np.diff(cost) <= 1e-9 * cost[0])
    chisq = sum(float(np.sum(np.abs((fr["vis"] - oracle.nufft_forward(fr["u"], fr["v"], x[k])) / fr["std"]) ** 2))
                for k, fr in enumerate(frames))
    assert abs(terms[0] - chisq) <= 1e-8 * chisq
    assert abs(terms[1] - x.sum()) <= 1e-10 * x.sum()
    assert abs(terms[2] - sum(mr.tsv(x[k]) for k in range(nt))) <= 1e-9 * terms[2]
    assert abs(terms[3] - mr.d_X(x, w)) <= 1e-9 * max(terms[3], 1e-300)
    for k in range(nt):
        sel = D["frame"] == k + 1
        mv = oracle.nufft_forward(D["u"][sel], D["v"][sel], x[k])
        assert np.abs(outvis[sel] - mv).max() <= 1e-8 * np.abs(mv).max()


def test_movie_without_temporal_term_is_independent_frames():
    """lambda_t = 0 decouples the frames: each converged frame is that frame's
    single-frame optimum (the GPU single-frame solve, same preset)."""
    nt, N = 3, 32
    D = _movie(nt, N, 150, 22)
    lam, lam2 = 0.05, 0.5
    print("case movie lambda_t = 0 vs independent single-frame solves", flush=True)
    x, cost, terms, outvis, info = sa.mfista_movie_nufft(D["vis"], D["u"], D["v"], D["std"], D["frame"], None, N, N,
                                                         None, lam, lam2, 0.0, nonneg_flag=1, cinit=1e3,
                                                         maxiter=30000, eps=0.0, rel_tol=1e-11)
    for k, fr in enumerate(_frames_of(D, nt)):
        xs, rs, _ = sa.mfista_imaging_core_nufft(fr["u"], fr["v"], fr["vis"].real, fr["vis"].imag, fr["std"], N, N,
                                                 maxiter=30000, eps=0.0, lambda_l1=lam, lambda_tsv=lam2, cinit=1e3,
                                                 nonneg_flag=1, rel_tol=1e-11)
        fk = mr.objective([fr], x[k:k + 1], lam, lam2, 0.0, np.zeros(0))
        e = np.linalg.norm(x[k] - xs) / np.linalg.norm(xs)
        print(f"  frame {k}: objective rel {abs(fk - rs['finalcost']) / rs['finalcost']:.2e} image relL2 {e:.2e}",
              flush=True)
        assert abs(fk - rs["finalcost"]) <= 1e-6 * rs["finalcost"]
        assert e <= 1e-3


def test_movie_box_and_errors():
    nt, N = 2, 32
    D = _movie(nt, N, 120, 23)
    box = np.ones((N, N), np.float32)
    box[:, :5] = 0
    x, cost, terms, outvis, info = sa.mfista_movie_nufft(D["vis"], D["u"], D["v"], D["std"], D["frame"], box, N, N,
                                                         np.full((N, N), 0.1), 0.05, 0.5, 1.0, nonneg_flag=1,
                                                         cinit=1e3, maxiter=300, eps=0.0)
    assert np.all(x[:, :, :5] == 0) and info["iterations"] == 300
    bad = D["frame"].copy()
    bad[bad == 2] = 3  # frame 2 left empty
    with pytest.raises(sa.MfistaError):
        sa.mfista_movie_nufft(D["vis"], D["u"], D["v"], D["std"], bad, None, N, N, None, 0.05, 0.5, 1.0,
                              maxiter=10)
