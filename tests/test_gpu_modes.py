"""GPU FFT mode, DFT mode, L1+TV (FGP) and the reference's own data files,
each against the FP64 oracle on the same inputs.

Bars: FP64 modes (FFT, DFT) image relative L2 <= 1e-6 and objective <= 1e-9
relative (same arithmetic, only summation order differs); FP32 NUFFT modes
image relative L2 <= 1e-3 and objective <= 1e-6 relative."""
import os

import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")


def _cmp(tag, xg, rg, o, img_tol, cost_tol):
    xo, ro = o["x"], o["result"]
    l2 = np.linalg.norm(xg - xo) / max(np.linalg.norm(xo), 1e-300)
    dc = abs(rg["finalcost"] - ro["finalcost"]) / abs(ro["finalcost"])
    print(f"  {tag}: ITER gpu={rg['ITER']} oracle={ro['ITER']} image relL2={l2:.3e} cost rel={dc:.3e}", flush=True)
    assert l2 <= img_tol, (tag, l2)
    assert dc <= cost_tol, (tag, dc)


def test_fft_mode_cfg1_64():
    P = synth.make_fft_problem(N=64, ncell=1500, seed=1001)
    kw = dict(maxiter=1000, eps=1e-5, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4)
    print("case fft-mode 64x64 M=3000 L1+TSV", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_fft(P["u_idx"], P["v_idx"], P["y_r"], P["y_i"], P["noise"], 64, 64,
                                           nonneg_flag=1, **kw)
    o = oracle.imaging_core_fft(P["u_idx"], P["v_idx"], P["y_r"], P["y_i"], P["noise"], 64, 64, kw["maxiter"],
                                kw["eps"], kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=1)
    _cmp("fft cfg1", xg, rg, o, 1e-6, 1e-9)
    assert rg["M"] == P["M"] // 2


def test_fft_mode_reference_data_tv():
    d = np.load(os.path.join(G, "ref_fft_data.npz"))
    N = int(d["NX"])
    print(f"case fft-mode reference fft_data.txt {N}x{N} L1+TV", flush=True)
    kw = dict(maxiter=60, eps=1e-5, lambda_l1=10.0, lambda_tv=50.0, lambda_tsv=0.0, cinit=1e4)
    xg, rg, _ = sa.mfista_imaging_core_fft(d["u"], d["v"], d["y_r"], d["y_i"], d["noise"], N, N, nonneg_flag=1,
                                           fgp_iter=30, **kw)
    o = oracle.imaging_core_fft(d["u"], d["v"], d["y_r"], d["y_i"], d["noise"], N, N, kw["maxiter"], kw["eps"],
                                kw["lambda_l1"], kw["lambda_tv"], 0.0, kw["cinit"], nonneg=1, fgp_iter=30)
    _cmp("fft ref-data TV", xg, rg, o, 1e-6, 1e-9)


@pytest.mark.parametrize("tv", [False, True])
def test_dft_mode(tv):
    rng = np.random.default_rng(5)
    NX = NY = 16
    N, M = NX * NY, 400
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    y = A @ np.maximum(rng.standard_normal(N), 0) + 0.05 * rng.standard_normal(M)
    kw = dict(maxiter=400, eps=1e-9, lambda_l1=0.05, lambda_tv=0.1 if tv else 0.0,
              lambda_tsv=0.0 if tv else 0.2, cinit=1.0)
    print(f"case dft-mode M={M} N={N} tv={tv}", flush=True)
    xg, rg, _, resid = sa.mfista_imaging_core_dft(y, A, NX, NY, nonneg_flag=1, fgp_iter=40, **kw)
    o = oracle.imaging_core_dft(y, A, NX, NY, kw["maxiter"], kw["eps"], kw["lambda_l1"], kw["lambda_tv"],
                                kw["lambda_tsv"], kw["cinit"], nonneg=1, fgp_iter=40)
    _cmp(f"dft tv={tv}", xg, rg, o, 1e-6, 1e-9)
    assert np.abs(resid - o["residual"]).max() <= 1e-6 * np.abs(o["residual"]).max()
    assert rg["M"] == M and rg["N"] == N


def test_nufft_tv_c24():
    P = synth.make_nufft_problem(N=64, M=2500, seed=1003, amax=0.35, kind="tv", coverage="disc")
    kw = dict(maxiter=3000, eps=0.0, lambda_l1=0.5, lambda_tv=5.0, lambda_tsv=0.0, cinit=1e4, rel_tol=1e-8)
    print("case nufft L1+TV C-24 64x64 M=2500 fgp=50", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50, **kw)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], kw["lambda_tv"], 0.0, kw["cinit"], nonneg=1,
                                  preset=oracle.PRESET_C24, fgp_iter=50, rel_tol=kw["rel_tol"])
    _cmp("nufft TV", xg, rg, o, 1e-3, 1e-6)


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_nufft_reference_data(preset):
    """c/nufft_data.txt (200x200, M=7648) with the author's run parameters
    (matlab/testdata4nufft.mat: lambda_l1=100, lambda_tsv=100, cinit=5e8,
    nonneg), run to a 1e-10 relative tolerance and compared with the optimum
    of the same problem (ref_nufft_solution.npz: L-BFGS-B on the oracle's FP64
    operators, pinned on the CPU by its KKT conditions).

    The objective of this problem is very flat: the oracle's own MFISTA at a
    1e-8 tolerance stops 7.6e-6 (CXX-12) and 2.1e-6 (C-24) above the optimum
    with images 5.4e-2 and 2.6e-2 away from it.  The GPU solve must reach the
    optimum's objective to 1e-6 (measured ~2e-7) and be at least 2x closer to
    the optimal image than that (bar 1e-2; measured 6e-3 .. 9e-3)."""
    d = np.load(os.path.join(G, "ref_nufft_data.npz"))
    sol = np.load(os.path.join(G, "ref_nufft_solution.npz"))
    N = int(d["NX"])
    kw = dict(maxiter=50000, eps=0.0, lambda_l1=100.0, lambda_tv=0.0, lambda_tsv=100.0, cinit=5e8, rel_tol=1e-10)
    print(f"case reference nufft_data.txt {N}x{N} M={d['u'].size} preset={preset}", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_nufft(d["u"], d["v"], d["vis_r"], d["vis_i"], d["std"], N, N,
                                             nonneg_flag=1, preset=preset, **kw)
    xo = sol[f"x{preset}"].astype(np.float64)
    fo = float(sol[f"finalcost{preset}"])
    l2 = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    dc = (rg["finalcost"] - fo) / abs(fo)
    print(f"  nufft ref-data: ITER gpu={rg['ITER']} image relL2={l2:.3e} cost (gpu - optimum)/optimum={dc:.3e}",
          flush=True)
    assert abs(dc) <= 1e-6, dc
    assert l2 <= 1e-2, l2


def test_clean_box_and_nonneg_flag():
    P = synth.make_nufft_problem(N=64, M=1500, seed=8, amax=0.35)
    box = np.ones((64, 64), np.float32)
    box[:, :20] = 0
    kw = dict(maxiter=400, eps=1e-5, lambda_l1=1.0, lambda_tsv=10.0, cinit=1e4)
    xg, rg, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=0, cl_box=box, preset=sa.PRESET_CXX12, **kw)
    assert np.all(xg[:, :20] == 0)
    x2, r2, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=2, preset=sa.PRESET_CXX12, **kw)
    assert r2["ITER"] == 0  # c/mfista_nufft_lib.c:634-637


def test_nufft_tv_512_m200000_against_oracle():
    """The 512^2 L1+TV workload (20 Gaussians + 2 rings, M = 200000 points in a
    disc of radius 0.48 cycles/px with density ~ 1/r, FGP 50 iterations,
    lambda_l1 = 0.5, lambda_tv = 5, nonneg, 24-tap preset) at full size.  The
    FP64 oracle needs ~4 s per iteration here, so both run the same first 8
    MFISTA iterations (no stopping rule involved) and are compared there; the
    full 500-iteration GPU solve must then lower the objective further."""
    P = synth.make_nufft_problem(N=512, M=200000, seed=1003, kind="tv", coverage="disc", amax=0.48)
    kw = dict(eps=0.0, lambda_l1=0.5, lambda_tv=5.0, lambda_tsv=0.0, cinit=1e4)
    print("case nufft L1+TV 512x512 M=200000 fgp=50, 8 iterations", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 512, 512,
                                             maxiter=8, nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50, **kw)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 512, 512, 8, kw["eps"],
                                  kw["lambda_l1"], kw["lambda_tv"], 0.0, kw["cinit"], nonneg=1,
                                  preset=oracle.PRESET_C24, fgp_iter=50)
    _cmp("nufft TV 512 (8 it)", xg, rg, o, 1e-3, 1e-6)
    assert rg["ITER"] == o["result"]["ITER"] == 8
    print("case nufft L1+TV 512x512 M=200000 fgp=50, 500 iterations", flush=True)
    xf, rf, info = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 512, 512,
                                                maxiter=500, nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50,
                                                **kw)
    print(f"  500 iterations: {info['solve_ms']:.0f} ms, finalcost {rf['finalcost']:.6e} "
          f"(8 iterations: {rg['finalcost']:.6e})", flush=True)
    assert rf["finalcost"] < rg["finalcost"]
