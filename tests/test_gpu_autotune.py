"""The persistent solver's launch-configuration autotuner and the gridding
tiles it chooses among (no reference counterpart; the reference's schedule is
fixed, its nearest mechanism being configure_tile, c++/mfista_nufft_lib.cpp:
643-678).

* every tile kernel (16x16, 32x16, 32x32 cells; 1, 2 and 4 cells per thread),
  with one and two workgroups per CU and with paired or single trial rounds,
  solves NUFFT L1+TSV (both presets) and L1+TV (C-24) to the oracle's optimum:
  image relative L2 <= 1e-3, objective <= 1e-6 -- the configuration is forced
  through a loaded tune table, so the test also covers the table's lookup;
* a batch launch with a forced 32x32 tile against the oracle;
* the tuner on the benchmark workload (256^2, M = 20000, 12 taps): 12
  candidates timed twice each, the fastest kept; a solve with autotune=True
  then times nothing new, and matches the oracle at its stopping count;
* the tuner on the 6x6 lambda grid at 128^2 (M = 8000), which includes the
  wide engine among its candidates: the tuned batch against the oracle's
  solutions of two grid points and against the untuned batch, a fixed 3000
  iterations;
* the FGP block shapes against the oracle, and the tuner on an L1+TV
  problem.
"""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu

TILES = [(16, 16), (32, 16), (32, 32)]


def _force(tmp_path, NX, NY, preset, reg, M, nprob, tile, bpc=2, paired=1, engine=0, fgp_tw=0):
    f = tmp_path / "tune.txt"
    f.write_text(f"{NX} {NY} {preset} {reg} {int(np.log2(M))} {int(np.log2(nprob))} {tile[0]} {tile[1]} "
                 f"{bpc} {paired} {engine} {fgp_tw} 0\n")
    sa.tune_clear()
    assert sa.tune_load(str(f)) == 1


def _cmp(tag, x, r, xo, Fo, l2_bar=1e-3, f_bar=1e-6):
    l2 = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    dF = abs(r["finalcost"] - Fo) / abs(Fo)
    print(f"  {tag}: ITER {r['ITER']} image relL2 {l2:.2e} objective rel {dF:.2e}", flush=True)
    assert l2 <= l2_bar
    assert dF <= f_bar


@pytest.mark.parametrize("tile,bpc,paired", [((16, 16), 2, 1), ((32, 16), 2, 1), ((32, 32), 2, 1),
                                             ((32, 16), 1, 0), ((32, 32), 1, 1)])
@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_tiles_l1_tsv_match_oracle(tmp_path, preset, tile, bpc, paired):
    P = synth.make_nufft_problem(N=64, M=2000, seed=1234, amax=0.35)
    # a fixed 5000 iterations, past convergence (a 1e-9 rule stops near 3700):
    # at that rule the gridding loop's FP32-atomic objective noise can end a
    # solve 1000 iterations early, 1.6e-3 from the optimum in image
    kw = dict(maxiter=5000, eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4,
              nonneg_flag=1, rel_tol=0.0)
    print(f"case tile {tile} bpc {bpc} paired {paired} preset {preset} N=64 M=2000 L1+TSV", flush=True)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=1,
                                  preset=preset, rel_tol=kw["rel_tol"])
    _force(tmp_path, 64, 64, preset, 1, 2000, 1, tile, bpc, paired)
    runs = sa.tune_runs()
    x, r, info = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                              preset=preset, autotune=True, loop=sa.LOOP_GRIDDING, **kw)
    assert sa.tune_runs() == runs  # the loaded entry was used, nothing was timed
    assert info["engine_used"] == sa.ENGINE_PERSISTENT
    if paired == 0:
        assert info["trial_rounds"] == info["total_trials"]
    assert r["ITER"] == 5000
    _cmp("tile", x, r, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


@pytest.mark.parametrize("tile", [(32, 16), (32, 32)])
def test_tiles_l1_tv_match_oracle(tmp_path, tile):
    P = synth.make_nufft_problem(N=64, M=2500, seed=1003, amax=0.35, kind="tv", coverage="disc")
    kw = dict(maxiter=3000, eps=0.0, lambda_l1=0.5, lambda_tv=5.0, lambda_tsv=0.0, cinit=1e4, rel_tol=1e-8)
    print(f"case tile {tile} nufft L1+TV C-24 64x64 M=2500 fgp=50", flush=True)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], kw["lambda_tv"], 0.0, kw["cinit"], nonneg=1,
                                  preset=oracle.PRESET_C24, fgp_iter=50, rel_tol=kw["rel_tol"])
    _force(tmp_path, 64, 64, sa.PRESET_C24, 2, 2500, 1, tile, 2, 0)
    x, r, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                           nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50, autotune=True, **kw)
    _cmp("tv tile", x, r, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


@pytest.mark.parametrize("fgp_tw", [64, 32, 16])
def test_fgp_block_shapes_match_oracle(tmp_path, fgp_tw):
    """The FGP iterations with pixels dealt in tiles of fgp_tw x (256 / fgp_tw)
    (the FGP block shapes the tuner chooses among) solve L1+TV to the
    oracle's optimum."""
    P = synth.make_nufft_problem(N=64, M=2500, seed=1003, amax=0.35, kind="tv", coverage="disc")
    kw = dict(maxiter=3000, eps=0.0, lambda_l1=0.5, lambda_tv=5.0, lambda_tsv=0.0, cinit=1e4, rel_tol=1e-8)
    print(f"case FGP block shape {fgp_tw}x{256 // fgp_tw} nufft L1+TV C-24 64x64 M=2500 fgp=50", flush=True)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], kw["lambda_tv"], 0.0, kw["cinit"], nonneg=1,
                                  preset=oracle.PRESET_C24, fgp_iter=50, rel_tol=kw["rel_tol"])
    _force(tmp_path, 64, 64, sa.PRESET_C24, 2, 2500, 1, (16, 16), 2, 0, fgp_tw=fgp_tw)
    x, r, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                           nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50, autotune=True, **kw)
    _cmp("fgp shape", x, r, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


def test_autotune_tv_shape():
    """The tuner on an L1+TV problem times 3 tiles x 2 workgroup counts x 4
    FGP block shapes; the tuned solve matches the oracle."""
    P = synth.make_nufft_problem(N=64, M=2500, seed=1003, amax=0.35, kind="tv", coverage="disc")
    args = (P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64)
    sa.tune_clear()
    print("case autotune nufft L1+TV C-24 64x64 M=2500 fgp=50", flush=True)
    best, cand = sa.autotune_nufft(*args, lambda_l1=0.5, lambda_tv=5.0, preset=sa.PRESET_C24, fgp_iter=50, iters=30)
    for c in cand:
        print(f"  tile {c['tile_x']}x{c['tile_y']} blocks/CU {c['blocks_per_cu']} fgp {c['fgp_tw']}: "
              f"{c['us_per_iter']:.1f} us/iteration", flush=True)
    print(f"  chosen: {best}", flush=True)
    assert len(cand) == 24 and {c["fgp_tw"] for c in cand} == {0, 16, 32, 64}
    kw = dict(maxiter=3000, eps=0.0, lambda_l1=0.5, lambda_tv=5.0, lambda_tsv=0.0, cinit=1e4, rel_tol=1e-8)
    x, r, _ = sa.mfista_imaging_core_nufft(*args, nonneg_flag=1, preset=sa.PRESET_C24, fgp_iter=50, autotune=True,
                                           **kw)
    o = oracle.imaging_core_nufft(*args, kw["maxiter"], kw["eps"], kw["lambda_l1"], kw["lambda_tv"], 0.0,
                                  kw["cinit"], nonneg=1, preset=oracle.PRESET_C24, fgp_iter=50, rel_tol=kw["rel_tol"])
    _cmp("tuned TV vs oracle", x, r, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


def test_batch_with_forced_tile_matches_oracle(tmp_path):
    """A batch launch with 32x32 gridding tiles: three lambda_tsv values over
    one 64^2 set, a fixed 3000 iterations (no stopping rule: the lambda_tsv =
    1 problem converges slowly, and rule-stopped runs end at points whose
    images differ by ~3e-3), each against the oracle run for as many."""
    P = synth.make_nufft_problem(N=64, M=1500, seed=1004, amax=0.35)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=64, NY=64,
                maxiter=3000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0, lambda_l1=1.0)
    probs = [dict(base, lambda_tsv=t) for t in (1.0, 10.0, 100.0)]
    print("case batch of 3 (64^2, M=1500) with 32x32 tiles, 3000 iterations, against the oracle", flush=True)
    _force(tmp_path, 64, 64, sa.PRESET_CXX12, 1, 1500, 3, (32, 32))
    xs, rs, info = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, autotune=True,
                                                      loop=sa.LOOP_GRIDDING)
    assert info["engine_used"] == sa.ENGINE_PERSISTENT
    for p, xb, rb in zip(probs, xs, rs):
        assert rb["ITER"] == 3000
        o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, 3000, 0.0, 1.0, 0.0,
                                      p["lambda_tsv"], 1e4, nonneg=1, preset=oracle.PRESET_CXX12, rel_tol=0.0)
        _cmp(f"batch tile 32x32, lambda_tsv {p['lambda_tsv']:g}", xb, rb, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


def test_autotune_benchmark_workload():
    P = synth.make_nufft_problem(N=256, M=20000, seed=1002)
    args = (P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 256, 256)
    sa.tune_clear()
    runs = sa.tune_runs()
    print("case autotune 256^2 M=20000 12 taps L1+TSV, 30 iterations per candidate", flush=True)
    best, cand = sa.autotune_nufft(*args, lambda_l1=1.0, lambda_tsv=100.0, preset=sa.PRESET_CXX12, iters=30)
    for c in cand:
        print(f"  loop {c['loop']} tile {c['tile_x']}x{c['tile_y']} blocks/CU {c['blocks_per_cu']} "
              f"paired {c['paired']}: {c['us_per_iter']:.1f} us/iteration", flush=True)
    print(f"  chosen: {best}", flush=True)
    # 12 gridding-loop candidates and the Toeplitz loop at 1 and 2 workgroups per CU
    assert len(cand) == 14 and sa.tune_runs() - runs == 28
    assert sorted(c["loop"] for c in cand) == [sa.LOOP_GRIDDING] * 12 + [sa.LOOP_TOEPLITZ] * 2
    assert best["us_per_iter"] == min(c["us_per_iter"] for c in cand)
    assert (best["NX"], best["NY"], best["preset"], best["reg"], best["m_log2"]) == (256, 256, sa.PRESET_CXX12, 1, 14)
    assert len(sa.tune_table()) == 1
    runs = sa.tune_runs()
    x, r, info = sa.mfista_imaging_core_nufft(*args, maxiter=50000, eps=0.0, lambda_l1=1.0, lambda_tsv=100.0,
                                              cinit=1e4, preset=sa.PRESET_CXX12, rel_tol=1e-6, autotune=True)
    assert sa.tune_runs() == runs
    assert info["loop_used"] == best["loop"]  # the tuned solve runs the chosen loop
    print(f"  tuned solve: ITER {r['ITER']} {info['kernel_ms']:.1f} ms, loop {info['loop_used']}", flush=True)
    o = oracle.imaging_core_nufft(*args[:5], 256, 256, r["ITER"], 0.0, 1.0, 0.0, 100.0, 1e4, nonneg=1,
                                  preset=oracle.PRESET_CXX12, rel_tol=0.0)
    _cmp("tuned vs oracle at the same iteration count", x, r, o["x"], o["result"]["finalcost"])
    sa.tune_clear()


def test_autotune_lambda_grid_batch():
    """The tuner on the 6x6 lambda grid at 128^2 (18 candidates: 12 persistent
    launches of the gridding loop, the Toeplitz loop at 1 and 2 workgroups per
    CU, the wide engine's gridding steps with 3 tiles and its Toeplitz steps);
    the tuned batch at a fixed
    3000 iterations against the oracle's solutions of two grid points
    (tests/golden/batch_oracle.npz) to the parity bars, and every problem
    against the default persistent batch (the two engines' trajectories differ
    by rounding: objective within 1e-5, image within 1e-2)."""
    import os

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "batch_oracle.npz"))
    P = synth.make_nufft_problem(N=128, M=8000, seed=1004)
    base = dict(u_dx=P["u"], v_dy=P["v"], vis_r=P["vis_r"], vis_i=P["vis_i"], vis_std=P["std"], NX=128, NY=128,
                maxiter=3000, eps=0.0, cinit=1e4, nonneg_flag=1, lambda_tv=0.0)
    probs = [dict(base, lambda_l1=10.0 ** a, lambda_tsv=10.0 ** b)
             for a in np.linspace(-2, 3, 6) for b in np.linspace(0, 5, 6)]
    sa.tune_clear()
    print("case autotune lambda grid 6x6 128^2 M=8000", flush=True)
    best, cand = sa.autotune_nufft_batch(probs, preset=sa.PRESET_CXX12, iters=30)
    for c in cand:
        print(f"  tile {c['tile_x']}x{c['tile_y']} blocks/CU {c['blocks_per_cu']} paired {c['paired']} "
              f"engine {c['engine']}: {c['us_per_iter']:.1f} us/iteration", flush=True)
    print(f"  chosen: {best}", flush=True)
    assert best["nprob_log2"] == 5 and len(cand) == 18
    assert sum(c["loop"] == sa.LOOP_TOEPLITZ for c in cand) == 3
    assert best["us_per_iter"] == min(c["us_per_iter"] for c in cand)
    xa, ra, ia = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12)
    xt, rt, it = sa.mfista_imaging_core_nufft_batch(probs, preset=sa.PRESET_CXX12, autotune=True)
    print(f"  3000 iterations: default {ia['kernel_ms']:.1f} ms, tuned {it['kernel_ms']:.1f} ms "
          f"(engine {it['engine_used']})", flush=True)
    for k in range(2):
        l1, ltsv = ref[f"grid{k}_lambda"]
        j = [i for i, p in enumerate(probs) if p["lambda_l1"] == l1 and p["lambda_tsv"] == ltsv][0]
        _cmp(f"tuned grid point ({l1:g}, {ltsv:g}) vs oracle", xt[j], rt[j], ref[f"grid{k}_x"],
             float(ref[f"grid{k}_F"]))
    for a, b, c, d in zip(xa, ra, xt, rt):
        assert b["ITER"] == d["ITER"] == 3000
        _cmp("tuned batch vs default batch", c, d, a, b["finalcost"], l2_bar=1e-2, f_bar=1e-5)
    sa.tune_clear()
