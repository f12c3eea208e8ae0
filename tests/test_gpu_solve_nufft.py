"""GPU MFISTA (NUFFT, L1+TSV) against the FP64 oracle, both run to convergence.

Parity bar: image relative L2 <= 1e-3 and final objective relative error
<= 1e-6 (FP32 operators, FP64 objective/backtracking and RESULT pass).
"""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu


def _problem(N, M, seed, amax):
    return synth.make_nufft_problem(N=N, M=M, seed=seed, amax=amax)


@pytest.mark.parametrize("engine", [sa.ENGINE_PERSISTENT, sa.ENGINE_STREAM])
@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_solve_l1_tsv_matches_oracle(preset, engine):
    # a fixed 5000 iterations, past convergence (a 1e-9 rule stops near 3700):
    # at that rule the FP32 gridding's objective noise can end a solve early
    P = _problem(64, 2000, 1234, 0.35)
    kw = dict(maxiter=5000, eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=10.0, cinit=1e4,
              nonneg_flag=1, rel_tol=0.0)
    print(f"case preset={preset} engine={engine} N=64 M=2000", flush=True)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"],
                                  nonneg=1, preset=preset, rel_tol=kw["rel_tol"])
    xg, rg, info = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                                preset=preset, engine=engine, **kw)
    xo, ro = o["x"], o["result"]
    l2 = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    dcost = abs(rg["finalcost"] - ro["finalcost"]) / abs(ro["finalcost"])
    print(f"  ITER gpu={rg['ITER']} oracle={ro['ITER']}  image relL2={l2:.3e}  cost rel={dcost:.3e}"
          f"  cost gpu={rg['finalcost']:.10e} oracle={ro['finalcost']:.10e}")
    assert rg["ITER"] == 5000
    assert l2 <= 1e-3
    assert dcost <= 1e-6
    assert rg["M"] == (2000 if preset == sa.PRESET_CXX12 else 1000)
    assert rg["N"] == 64 * 64 and rg["NX"] == 64


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_engines_agree(preset):
    """The persistent cooperative launch (FP64 iterates, visibility-domain
    momentum) and the kernel-per-stage stream (FP32 iterates, FP64 forward at
    the momentum point) are independent implementations of the same loop:
    run to a 1e-9 relative tolerance on a second problem size they reach the
    same optimum (mid-run trajectories differ: the backtracking decisions see
    different rounding)."""
    P = _problem(96, 3000, 555, 0.35)
    kw = dict(maxiter=20000, eps=0.0, lambda_l1=1.0, lambda_tsv=20.0, cinit=1e4, nonneg_flag=1, rel_tol=1e-9)
    print(f"case engines preset={preset} N=96 M=3000 converged", flush=True)
    xa, ra, ia = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 96, 96,
                                              preset=preset, engine=sa.ENGINE_PERSISTENT, **kw)
    xb, rb, ib = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 96, 96,
                                              preset=preset, engine=sa.ENGINE_STREAM, **kw)
    l2 = np.linalg.norm(xa - xb) / np.linalg.norm(xb)
    dc = abs(ra["finalcost"] - rb["finalcost"]) / rb["finalcost"]
    print(f"  ITER persistent={ra['ITER']} stream={rb['ITER']} relL2={l2:.2e} cost rel={dc:.2e} "
          f"ms persistent={ia['solve_ms']:.1f} stream={ib['solve_ms']:.1f}")
    assert l2 <= 1e-3 and dc <= 1e-6


@pytest.mark.parametrize("preset", [sa.PRESET_CXX12, sa.PRESET_C24])
def test_paired_trial_rounds_match_single_trials(preset):
    """The persistent engine evaluates the backtracking trials c and c*ETA of
    an L1+TSV iteration together and takes them in order; with one trial per
    round it runs the reference's loop trial by trial.  Mid-run the two cannot
    be compared bit for bit: repeated runs of either mode already differ by
    ~1e-4 after 300 iterations (FP32 gridding atomics, amplified by the
    momentum).  A relative stopping rule is no fair end point either: at 1e-9,
    about half of the GPU runs of this problem stop on a plateau at ~4600
    iterations instead of ~5540 (objective 8e-8 above the optimum, image
    1.2e-3 away), in either mode.  So both modes and the FP64 oracle run the
    same 6000 iterations, past convergence, and must agree there."""
    P = _problem(64, 2000, 321, 0.35)
    kw = dict(maxiter=6000, eps=0.0, lambda_l1=1.0, lambda_tsv=100.0, cinit=1e4, nonneg_flag=1)
    print(f"case paired vs single trial rounds preset={preset} N=64 M=2000 lambda_tsv=100, 6000 iterations",
          flush=True)
    xa, ra, ia = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                              preset=preset, **kw)
    xb, rb, ib = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                              preset=preset, single_trials=True, **kw)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"],
                                  kw["eps"], kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=1,
                                  preset=preset)
    fo, xo = o["result"]["finalcost"], o["x"]
    dc = abs(ra["finalcost"] - rb["finalcost"]) / rb["finalcost"]
    print(f"  paired: ITER {ra['ITER']} trials {ia['total_trials']} rounds {ia['trial_rounds']} "
          f"{ia['kernel_ms']:.1f} ms; single: ITER {rb['ITER']} trials {ib['total_trials']} rounds "
          f"{ib['trial_rounds']} {ib['kernel_ms']:.1f} ms; mutual cost rel {dc:.2e}", flush=True)
    for name, x, r in (("paired", xa, ra), ("single", xb, rb)):
        l2 = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        c = abs(r["finalcost"] - fo) / fo
        print(f"  {name} vs oracle: image relL2 {l2:.2e} cost rel {c:.2e}", flush=True)
        assert l2 <= 1e-3 and c <= 1e-6
    assert ra["ITER"] == rb["ITER"] == 6000
    assert ib["trial_rounds"] == ib["total_trials"]
    assert ia["trial_rounds"] < ia["total_trials"]
    assert dc <= 1e-7


def test_result_fields_consistent():
    P = _problem(64, 1500, 99, 0.35)
    x, r, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                           maxiter=300, eps=1e-5, lambda_l1=1.0, lambda_tsv=10.0,
                                           cinit=1e4, nonneg_flag=1, preset=sa.PRESET_CXX12)
    ro = oracle.result_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], x, 1.0, 0.0, 10.0,
                             preset=oracle.PRESET_CXX12)
    for k in ("sq_error", "l1cost", "tsvcost", "finalcost"):
        assert abs(r[k] - ro[k]) <= 1e-9 * abs(ro[k]), k
    assert r["N_active"] == ro["N_active"]
    assert (x >= 0).all()


def test_both_regularisers_rejected():
    P = _problem(64, 500, 3, 0.35)
    with pytest.raises(sa.MfistaError) as e:
        sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                     maxiter=10, lambda_tv=1.0, lambda_tsv=1.0)
    assert e.value.status == -1


@pytest.mark.parametrize("loop", [sa.LOOP_GRIDDING, sa.LOOP_AUTO])
def test_persistent_trace_summary(tmp_path, monkeypatch, loop):
    """MFISTA_PERSISTENT_TRACE=<file>: the persistent solver timestamps every
    grid barrier; the summary's stage and tail times add up to the launch
    (the gridding loop: five stages per iteration; the Toeplitz loop, the
    default here: three per trial round)."""
    P = _problem(64, 2000, 7, 0.35)
    path = tmp_path / "trace.txt"
    monkeypatch.setenv("MFISTA_PERSISTENT_TRACE", str(path))
    x, r, info = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                              maxiter=300, eps=0.0, lambda_l1=1.0, lambda_tsv=10.0, cinit=1e4,
                                              nonneg_flag=1, preset=sa.PRESET_CXX12, loop=loop)
    text = path.read_text()
    print(text, flush=True)
    counts = {ln.split()[0]: int(ln.split()[1]) for ln in text.splitlines() if ln.startswith(("P1", "P2", "T"))}
    if loop == sa.LOOP_GRIDDING:
        assert info["loop_used"] == sa.LOOP_GRIDDING and "gridding loop" in text
        for stage in ("P1-grid", "P2-colinv", "T1-rows", "T2-cols", "T3-degrid"):
            assert stage in text
        assert counts["P1-grid"] == 300 and counts["P2-colinv"] == 300 and counts["T1-rows"] == info["trial_rounds"]
    else:
        assert info["loop_used"] == sa.LOOP_TOEPLITZ and "Toeplitz loop" in text
        for stage in ("T1-rows", "T2-conv", "T3-c2r"):
            assert counts[stage] == info["trial_rounds"]
    total = float(text.split("sum over traced barriers")[1].split()[0])
    assert 0.7 * info["kernel_ms"] <= total <= 1.05 * info["kernel_ms"]
