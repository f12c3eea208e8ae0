"""GPU paths that the default workloads leave out, each against the FP64
oracle on the same inputs:

  * the signed L1 prox, nonneg_flag = 0 (c/mfista_tools.c:81-112), in the
    NUFFT L1+TSV loop of both presets and engines, and in FFT mode;
  * L1+TV through FGP_L1 (c/mfista_TV_lib.c:124-253, nonneg_flag = 0) and
    L1+TV with a CLEAN box (FGP_nonneg_box, :256-384), both engines, and in
    DFT mode;
  * the reference's own void entry points mfista_imaging_core_{nufft,fft,dft}
    (c/mfista.h:151-178) called through the C ABI with the reference's
    defaults: the 24-tap preset and the absolute stopping rule
    cost[t-50] - cost[t] < eps sum|V|^2 / M (c/mfista_nufft_lib.c:741,
    923-926), the RESULT fields, and the early return when both lambda_tv and
    lambda_tsv are positive (c/mfista_nufft_lib.c:938-941).

Bars: FP32 NUFFT image relative L2 <= 1e-3 and objective <= 1e-6 relative;
FP64 FFT/DFT modes 1e-6 and 1e-9.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu


def _signed_problem(N, M, seed):
    """A sky with negative regions (so the signed prox matters), 12 arcs."""
    rng = np.random.default_rng(seed)
    x = synth.gaussian_blobs(N, rng, 3) - 0.7 * synth.gaussian_blobs(N, rng, 2)
    u, v = synth.elliptical_arcs(M, rng, amax=0.35)
    vis = synth.dft_vis(x, u, v)
    sig = 0.01 * np.abs(vis).max()
    vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    return {"u": u, "v": v, "vis_r": vis.real.copy(), "vis_i": vis.imag.copy(), "std": np.full(M, sig)}


def _cmp(tag, xg, rg, o, img_tol, cost_tol):
    xo, ro = o["x"], o["result"]
    l2 = np.linalg.norm(xg - xo) / max(np.linalg.norm(xo), 1e-300)
    dc = abs(rg["finalcost"] - ro["finalcost"]) / abs(ro["finalcost"])
    print(f"  {tag}: ITER gpu={rg['ITER']} oracle={ro['ITER']} image relL2={l2:.3e} cost rel={dc:.3e}", flush=True)
    assert l2 <= img_tol, (tag, l2)
    assert dc <= cost_tol, (tag, dc)


@pytest.mark.parametrize("engine", [sa.ENGINE_PERSISTENT, sa.ENGINE_STREAM])
@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_nufft_signed_l1_tsv(preset, engine):
    P = _signed_problem(64, 2000, 71)
    kw = dict(maxiter=20000, eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4, rel_tol=1e-9)
    print(f"case nufft signed L1+TSV preset={preset} engine={engine} 64x64 M=2000", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=0, preset=preset, engine=engine, **kw)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"], 0.0,
                                  kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=0, preset=preset,
                                  rel_tol=kw["rel_tol"])
    assert (o["x"] < 0).any() and (xg < 0).any()
    _cmp("signed L1+TSV", xg, rg, o, 1e-3, 1e-6)


@pytest.mark.parametrize("engine", [sa.ENGINE_PERSISTENT, sa.ENGINE_STREAM])
@pytest.mark.parametrize("case", ["fgp_l1", "box"])
def test_nufft_tv_signed_and_box(case, engine):
    """L1+TV (24-tap preset, FGP 50 iterations): nonneg_flag = 0 runs FGP_L1;
    nonneg_flag = 1 with a CLEAN box runs FGP_nonneg_box.  The signed TV
    problem converges slowly (a stopping rule ends the GPU and the oracle
    solves at different points whose images differ by ~2e-3), so both run a
    fixed 20000 iterations; the oracle's solutions are the fixtures of
    tests/golden/tv_oracle.npz (make_tv_reference.py, same inputs)."""
    import hashlib
    import os

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "tv_oracle.npz"))
    if case == "fgp_l1":
        P = _signed_problem(64, 2500, 72)
        nonneg, box = 0, None
    else:
        P = synth.make_nufft_problem(N=64, M=2500, seed=73, amax=0.35, kind="tv", coverage="disc")
        nonneg = 1
        box = np.ones((64, 64), np.float32)
        box[:, :12] = 0.0
        box[50:, :] = 0.0
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    assert h.hexdigest() == str(ref[f"{case}_sha256"])
    kw = dict(maxiter=20000, eps=0.0, lambda_l1=0.5, lambda_tv=2.0, lambda_tsv=0.0, cinit=1e4, rel_tol=0.0)
    print(f"case nufft L1+TV {case} engine={engine} 64x64 M=2500 fgp=50, 20000 iterations", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=nonneg, cl_box=box, preset=sa.PRESET_C24, fgp_iter=50,
                                             engine=engine, **kw)
    o = {"x": ref[f"{case}_x"], "result": {"finalcost": float(ref[f"{case}_F"]), "ITER": 20000}}
    if box is not None:
        assert np.all(xg[box == 0] == 0)
    else:
        assert (o["x"] < 0).any() and (xg < 0).any()
    _cmp(f"TV {case}", xg, rg, o, 1e-3, 1e-6)


@pytest.mark.parametrize("case", ["signed", "box"])
def test_dft_tv_signed_and_box(case):
    rng = np.random.default_rng(81)
    NX = NY = 12
    N, M = NX * NY, 300
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    xs = rng.standard_normal(N) if case == "signed" else np.maximum(rng.standard_normal(N), 0)
    y = A @ xs + 0.05 * rng.standard_normal(M)
    box = None
    if case == "box":
        box = np.ones(N, np.float32)
        box[::5] = 0.0
    nonneg = 0 if case == "signed" else 1
    kw = dict(maxiter=500, eps=1e-12, lambda_l1=0.05, lambda_tv=0.1, lambda_tsv=0.0, cinit=1.0)
    print(f"case dft L1+TV {case} M={M} N={N}", flush=True)
    xg, rg, _, _ = sa.mfista_imaging_core_dft(y, A, NX, NY, nonneg_flag=nonneg, cl_box=box, fgp_iter=40, **kw)
    o = oracle.imaging_core_dft(y, A, NX, NY, kw["maxiter"], kw["eps"], kw["lambda_l1"], kw["lambda_tv"], 0.0,
                                kw["cinit"], nonneg=nonneg, cl_box=box, fgp_iter=40)
    _cmp(f"dft TV {case}", xg, rg, o, 1e-6, 1e-9)


def test_fft_mode_signed_l1_tsv():
    P = synth.make_fft_problem(N=64, ncell=1500, seed=82)
    kw = dict(maxiter=600, eps=1e-9, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4)
    print("case fft-mode signed L1+TSV 64x64", flush=True)
    xg, rg, _ = sa.mfista_imaging_core_fft(P["u_idx"], P["v_idx"], P["y_r"], P["y_i"], P["noise"], 64, 64,
                                           nonneg_flag=0, **kw)
    o = oracle.imaging_core_fft(P["u_idx"], P["v_idx"], P["y_r"], P["y_i"], P["noise"], 64, 64, kw["maxiter"],
                                kw["eps"], kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=0)
    _cmp("fft signed", xg, rg, o, 1e-6, 1e-9)


# ---------------------------------------------------------------- reference void entry points
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _sentinel():
    r = sa.RESULT()
    r.ITER, r.M, r.finalcost, r.sq_error = -7, -7, -7.0, -7.0
    return r


def test_reference_nufft_entry_defaults():
    """mfista_imaging_core_nufft as the reference's drivers call it: 24-tap
    preset, absolute rule with eps = 1e-6 (scaled by mean |V|^2)."""
    L = sa.lib()
    P = synth.make_nufft_problem(N=64, M=2000, seed=83, amax=0.35)
    N, M, eps, maxiter = 64, 2000, 1e-6, 20000
    u, v, vr, vi, sd = (np.ascontiguousarray(P[k]) for k in ("u", "v", "vis_r", "vis_i", "std"))
    x0, xo = np.zeros(N * N), np.zeros(N * N)
    res = sa.RESULT()
    print("case reference entry mfista_imaging_core_nufft 64x64 M=2000 eps=1e-6", flush=True)
    L.mfista_imaging_core_nufft(_ptr(u), _ptr(v), _ptr(vr), _ptr(vi), _ptr(sd), M, N, N, maxiter, eps, 1.0, 0.0,
                                10.0, 1e4, _ptr(x0), _ptr(xo), 1, 0, None, C.byref(res))
    o = oracle.imaging_core_nufft(u, v, vr, vi, sd, N, N, maxiter, eps, 1.0, 0.0, 10.0, 1e4, nonneg=1,
                                  preset=oracle.PRESET_C24)
    r, ro = res.as_dict(), o["result"]
    _cmp("reference entry", xo.reshape(N, N), r, o, 1e-3, 1e-6)
    band = 0.25 * ro["ITER"] + 50
    assert abs(r["ITER"] - ro["ITER"]) <= band, (r["ITER"], ro["ITER"])
    assert r["M"] == M // 2 and r["N"] == N * N and r["NX"] == N and r["NY"] == N  # c/mfista_nufft_lib.c:849
    assert r["maxiter"] == maxiter and r["nonneg"] == 1
    assert (r["lambda_l1"], r["lambda_tv"], r["lambda_tsv"]) == (1.0, 0.0, 10.0)
    # the two solves stop at different iterations of the same absolute rule:
    # the objective agrees to the parity bar, its terms to the image bar
    assert r["finalcost"] == pytest.approx(ro["finalcost"], rel=1e-6)
    for k in ("sq_error", "l1cost", "tsvcost"):
        assert r[k] == pytest.approx(ro[k], rel=1e-3), k
    assert r["mean_sq_error"] == pytest.approx(r["sq_error"] / M, rel=1e-12)
    assert r["finalcost"] == pytest.approx(r["sq_error"] / 2 + r["l1cost"] + 10.0 * r["tsvcost"], rel=1e-12)
    assert abs(r["N_active"] - ro["N_active"]) <= 0.02 * ro["N_active"] + 2
    assert r["comp_time"] > 0 and r["Lip_const"] > 0


def test_reference_entries_reject_both_regularisers(capfd):
    L = sa.lib()
    P = synth.make_nufft_problem(N=64, M=500, seed=84, amax=0.35)
    u, v, vr, vi, sd = (np.ascontiguousarray(P[k]) for k in ("u", "v", "vis_r", "vis_i", "std"))
    x0, xo = np.zeros(64 * 64), np.full(64 * 64, 3.0)
    res = _sentinel()
    L.mfista_imaging_core_nufft(_ptr(u), _ptr(v), _ptr(vr), _ptr(vi), _ptr(sd), 500, 64, 64, 100, 1e-5, 1.0, 1.0,
                                1.0, 1e4, _ptr(x0), _ptr(xo), 1, 0, None, C.byref(res))
    out = capfd.readouterr().out
    assert "You cannot set both of lambda_TV and lambda_TSV positive." in out
    assert res.ITER == -7 and res.finalcost == -7.0 and np.all(xo == 3.0)  # RESULT and xout untouched


def test_reference_fft_and_dft_entries():
    L = sa.lib()
    Pf = synth.make_fft_problem(N=32, ncell=300, seed=85)
    ui = np.ascontiguousarray(Pf["u_idx"], np.int32)
    vi_ = np.ascontiguousarray(Pf["v_idx"], np.int32)
    yr, yi, ns = (np.ascontiguousarray(Pf[k]) for k in ("y_r", "y_i", "noise"))
    x0, xo = np.zeros(32 * 32), np.zeros(32 * 32)
    res = sa.RESULT()
    print("case reference entry mfista_imaging_core_fft 32x32", flush=True)
    L.mfista_imaging_core_fft(ui.ctypes.data_as(_ip), vi_.ctypes.data_as(_ip), _ptr(yr), _ptr(yi), _ptr(ns),
                              ui.size, 32, 32, 2000, 1e-7, 1.0, 0.0, 5.0, 1e4, _ptr(x0), _ptr(xo), 1, 0, 0, None,
                              C.byref(res))
    o = oracle.imaging_core_fft(ui, vi_, yr, yi, ns, 32, 32, 2000, 1e-7, 1.0, 0.0, 5.0, 1e4, nonneg=1)
    _cmp("reference fft entry", xo.reshape(32, 32), res.as_dict(), o, 1e-6, 1e-9)
    assert res.ITER == o["result"]["ITER"] and res.M == ui.size // 2

    rng = np.random.default_rng(86)
    NX = NY = 8
    Nn, M = NX * NY, 120
    A = np.asfortranarray(rng.standard_normal((M, Nn)) / np.sqrt(M))
    y = A @ np.maximum(rng.standard_normal(Nn), 0) + 0.05 * rng.standard_normal(M)
    x0, xo = np.zeros(Nn), np.zeros(Nn)
    Mi, Ni = C.c_int(M), C.c_int(Nn)
    res = sa.RESULT()
    print("case reference entry mfista_imaging_core_dft M=120 N=64", flush=True)
    L.mfista_imaging_core_dft(_ptr(np.ascontiguousarray(y)), A.ctypes.data_as(_dp), C.byref(Mi), C.byref(Ni), NX,
                              NY, 5000, 1e-10, 0.05, 0.0, 0.2, 1.0, _ptr(x0), _ptr(xo), 1, 0, 0, None, C.byref(res))
    o = oracle.imaging_core_dft(y, A, NX, NY, 5000, 1e-10, 0.05, 0.0, 0.2, 1.0, nonneg=1)
    _cmp("reference dft entry", xo.reshape(NX, NY), res.as_dict(), o, 1e-6, 1e-9)
    assert res.residual  # malloc'ed with M doubles; the caller frees it (c/mfista_dft_lib.c:454)
    resid = np.ctypeslib.as_array(res.residual, shape=(M,)).copy()
    C.CDLL(None).free(C.cast(res.residual, C.c_void_p))
    assert np.abs(resid - o["residual"]).max() <= 1e-6 * np.abs(o["residual"]).max()
