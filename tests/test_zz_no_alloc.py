"""The allocation-free entry point mfista_nufft_solve_device_noalloc against
mfista_nufft_solve_device, called through the C ABI with device pointers.

No counterpart in the reference (the solve itself is
c/mfista_nufft_lib.c:909-957).  Both entry points run the same persistent
solve; in the deterministic mode (gridding without floating-point atomics)
their images, objectives and iteration counts must agree bit for bit, which
implies the parity bar (image relative L2 <= 1e-3, objective <= 1e-6).  Two
allocation-free calls in a row, the second on a smaller problem, so that the
second runs on scratch memory holding the first's contents; neither may make a
device allocation.  A problem too large for the scratch array is rejected
with MFISTA_E_SIZE, and the entry point works again afterwards.
"""
import ctypes as C

import numpy as np
import pytest

import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu

KW = dict(eps=0.0, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=30.0, cinit=1e4, nonneg_flag=1)


def _bind():
    L = sa.lib()
    f = L.mfista_nufft_solve_device_noalloc
    f.restype = C.c_int
    f.argtypes = L.mfista_nufft_solve_device.argtypes
    L.mfista_device_allocations.restype = C.c_long
    L.mfista_device_allocations.argtypes = []
    return L


def _tensors(P, N):
    import torch

    dev = torch.device("cuda:0")
    t = {k: torch.tensor(P[k], dtype=torch.float64, device=dev) for k in ("u", "v", "vis_r", "vis_i", "std")}
    t["x0"] = torch.zeros(N * N, dtype=torch.float64, device=dev)
    t["xo"] = torch.full((N * N,), -7.0, dtype=torch.float64, device=dev)
    return t


def _call(fn, t, N, maxiter):
    import torch

    res = sa.RESULT()
    o = sa.make_opts(sa.PRESET_CXX12, sa.FGPITER, 0.0, deterministic=True)
    stream = torch.cuda.current_stream().cuda_stream
    rc = fn(t["u"].data_ptr(), t["v"].data_ptr(), t["vis_r"].data_ptr(), t["vis_i"].data_ptr(), t["std"].data_ptr(),
            t["u"].numel(), N, N, maxiter, KW["eps"], KW["lambda_l1"], KW["lambda_tv"], KW["lambda_tsv"],
            KW["cinit"], t["x0"].data_ptr(), t["xo"].data_ptr(), KW["nonneg_flag"], 0, None, C.byref(res), C.byref(o),
            stream)
    return rc, res.as_dict(), t["xo"].cpu().numpy().copy()


def test_noalloc_matches_device_entry():
    L = _bind()
    cases = [(128, 8000, 31, 600), (64, 1500, 32, 400)]  # the second call is the smaller problem
    probs = [(N, synth.make_nufft_problem(N=N, M=M, seed=s, amax=0.35), it) for N, M, s, it in cases]
    refs = []
    for N, P, it in probs:
        print(f"case reference entry N={N} M={P['u'].size} {it} iterations", flush=True)
        rc, r, x = _call(L.mfista_nufft_solve_device, _tensors(P, N), N, it)
        assert rc == 0
        refs.append((r, x))
    n0 = L.mfista_device_allocations()
    for (N, P, it), (rr, xr) in zip(probs, refs):
        print(f"case allocation-free entry N={N} M={P['u'].size} {it} iterations", flush=True)
        rc, r, x = _call(L.mfista_nufft_solve_device_noalloc, _tensors(P, N), N, it)
        assert rc == 0
        l2 = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        dc = abs(r["finalcost"] - rr["finalcost"]) / rr["finalcost"]
        print(f"  ITER {r['ITER']} / {rr['ITER']}  image relL2 {l2:.2e}  objective rel {dc:.2e}", flush=True)
        assert r["ITER"] == rr["ITER"] == it
        assert l2 <= 1e-3 and dc <= 1e-6
        assert np.array_equal(x, xr) and r["finalcost"] == rr["finalcost"]
    assert L.mfista_device_allocations() == n0, "the allocation-free entry point allocated device memory"


def test_noalloc_rejects_oversize_then_recovers():
    L = _bind()
    rng = np.random.default_rng(5)
    N, M = 512, 300000
    big = {"u": rng.uniform(-1.0, 1.0, M), "v": rng.uniform(-1.0, 1.0, M), "vis_r": rng.standard_normal(M),
           "vis_i": rng.standard_normal(M), "std": np.ones(M)}
    print(f"case allocation-free oversize N={N} M={M}", flush=True)
    n0 = L.mfista_device_allocations()
    rc, _, _ = _call(L.mfista_nufft_solve_device_noalloc, _tensors(big, N), N, 10)
    assert rc == -4  # MFISTA_E_SIZE
    P = synth.make_nufft_problem(N=64, M=1500, seed=33, amax=0.35)
    print("case allocation-free after the rejection N=64 M=1500", flush=True)
    rc, r, x = _call(L.mfista_nufft_solve_device_noalloc, _tensors(P, 64), 64, 200)
    rc2, r2, x2 = _call(L.mfista_nufft_solve_device, _tensors(P, 64), 64, 200)
    assert rc == 0 and rc2 == 0 and r["ITER"] == 200
    assert np.array_equal(x, x2)
    assert L.mfista_device_allocations() >= n0
