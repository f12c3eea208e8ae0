"""Leave-one-out error estimate on the GPU (compute_LOOE_L1 / compute_LOOE_L1_TSV,
mfista_looe_dft) against the oracle (oracle/looe_ref.py, a restatement of
c/looe_lib.c pinned in test_oracle_looe.py) on the same inputs.

Bar: FP64 throughout; LOOE and its standard deviation within 1e-9 relative
(the Cholesky runs in a different order from the oracle's).  The
not-positive-definite and empty-active-set cases are exact."""
import ctypes as C

import numpy as np
import pytest

from oracle import looe_ref as lr
import sparseimaging_amd as sa

pytestmark = pytest.mark.gpu


def _case(M, NX, NY, frac, seed, lstsq=False):
    rng = np.random.default_rng(seed)
    N = NX * NY
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    x = np.where(rng.uniform(size=N) < frac, rng.uniform(0.5, 2.0, N), 0.0)
    y = A @ x + 0.05 * rng.standard_normal(M)
    if lstsq:  # least squares on the active set, so y - A x is orthogonal to it
        S = np.flatnonzero(x)
        x[S] = np.linalg.lstsq(A[:, S], y, rcond=None)[0]
    return A, x, y, y - A @ x


@pytest.mark.parametrize("M,NX,NY,frac,lt", [(200, 8, 8, 0.3, 0.0), (200, 8, 8, 0.3, 0.5),
                                             (3000, 32, 32, 0.25, 0.2), (500, 12, 10, 0.5, 1.0)])
def test_looe_matches_oracle(M, NX, NY, frac, lt):
    A, x, y, r = _case(M, NX, NY, frac, 11 + M + NX)
    print(f"case LOOE M={M} N={NX}x{NY} active~{frac} lambda_tsv={lt}", flush=True)
    g = sa.compute_looe(A, x, r, NX, NY, lt)
    o = lr.compute_looe(A, x, r, NX, NY, lt)
    print(f"  active {g['n_active']}: GPU {g['looe']:.12e} +- {g['looe_std']:.6e}, oracle {o[0]:.12e} +- "
          f"{o[2]:.6e}", flush=True)
    assert g["hessian_positive"] == 1 and g["n_active"] == int(np.count_nonzero(x))
    assert abs(g["looe"] - o[0]) <= 1e-9 * abs(o[0])
    assert abs(g["looe_std"] - o[2]) <= 1e-9 * abs(o[2])


def test_reference_signatures():
    """compute_LOOE_L1 / compute_LOOE_L1_TSV called as the reference declares
    them (c/looe_lib.c:195, :276): LOOE returned, looe_m and looe_std set."""
    A, x, y, r = _case(300, 10, 10, 0.3, 5)
    L = sa.lib()
    Af = np.asfortranarray(A)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    M, N = C.c_int(300), C.c_int(100)
    lm, ls = C.c_double(0), C.c_double(0)
    v = L.compute_LOOE_L1(C.byref(M), C.byref(N), 1.0, dp(y), dp(Af), dp(x), dp(r), C.byref(lm), C.byref(ls))
    o = lr.compute_looe(A, x, r)
    assert abs(v - o[0]) <= 1e-9 * o[0] and v == lm.value and abs(ls.value - o[2]) <= 1e-9 * o[2]
    v = L.compute_LOOE_L1_TSV(C.byref(M), C.byref(N), 10, 10, 1.0, 0.7, dp(y), dp(Af), dp(x), dp(r),
                              C.byref(lm), C.byref(ls))
    o = lr.compute_looe(A, x, r, 10, 10, 0.7)
    assert abs(v - o[0]) <= 1e-9 * o[0] and v == lm.value and abs(ls.value - o[2]) <= 1e-9 * o[2]


def test_not_positive_definite_and_empty_active_set():
    rng = np.random.default_rng(9)
    M, N = 20, 36  # 36 active pixels, 20 observations: A_s^T A_s is singular
    A = rng.standard_normal((M, N))
    x = np.ones(N)
    r = rng.standard_normal(M)
    print("case LOOE singular Hessian", flush=True)
    g = sa.compute_looe(A, x, r, 6, 6, 0.0)
    assert (g["looe"], g["looe_m"], g["looe_std"], g["hessian_positive"]) == (-1.0, 0.0, 0.0, 0)
    assert lr.compute_looe(A, x, r) == (-1.0, 0.0, 0.0)
    print("case LOOE empty active set", flush=True)
    g = sa.compute_looe(A, np.zeros(N), r)
    o = lr.compute_looe(A, np.zeros(N), r)
    assert g["n_active"] == 0 and g["hessian_positive"] == 1
    assert abs(g["looe"] - o[0]) <= 1e-14 * o[0] and abs(g["looe_std"] - o[2]) <= 1e-12 * o[2]


def test_hat_matrix_identity_on_gpu():
    """All pixels active, lambda_tsv = 0, x the least-squares fit: the GPU's
    (y - A x)_m / (1 - d_m) terms must reproduce the brute-force leave-one-out
    residuals (through their mean square / 2)."""
    rng = np.random.default_rng(4)
    M, N = 60, 16
    A = rng.standard_normal((M, N))
    y = A @ rng.standard_normal(N) + 0.1 * rng.standard_normal(M)
    x = np.linalg.lstsq(A, y, rcond=None)[0]
    loo = []
    for m in range(M):
        keep = np.arange(M) != m
        xm = np.linalg.lstsq(A[keep], y[keep], rcond=None)[0]
        loo.append((y[m] - A[m] @ xm) ** 2 / 2)
    g = sa.compute_looe(A, x, y - A @ x)
    assert abs(g["looe"] - np.mean(loo)) <= 1e-9 * np.mean(loo)
    assert abs(g["looe_std"] - np.std(loo, ddof=1)) <= 1e-8 * np.std(loo, ddof=1)


def test_looe_of_a_dft_solve():
    """The intended use (c/mfista_dft_lib.c:506-530): solve L1+TSV in DFT mode
    on the GPU, then estimate the LOOE of that solution."""
    rng = np.random.default_rng(21)
    NX = NY = 12
    M, N = 400, NX * NY
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    truth = np.zeros(N)
    truth[rng.choice(N, 15, replace=False)] = rng.uniform(1, 3, 15)
    y = A @ truth + 0.02 * rng.standard_normal(M)
    print("case LOOE after a DFT-mode L1+TSV solve", flush=True)
    xs, res, _, _ = sa.mfista_imaging_core_dft(y, A, NX, NY, maxiter=3000, eps=1e-9, lambda_l1=0.05,
                                               lambda_tsv=0.05, cinit=1.0, nonneg_flag=1)
    x = xs.ravel()
    r = y - A @ x
    g = sa.compute_looe(A, x, r, NX, NY, 0.05)
    o = lr.compute_looe(A, x, r, NX, NY, 0.05)
    print(f"  active {g['n_active']}: GPU {g['looe']:.10e}, oracle {o[0]:.10e}", flush=True)
    assert 0 < g["n_active"] < N and g["hessian_positive"] == 1
    assert abs(g["looe"] - o[0]) <= 1e-9 * o[0] and abs(g["looe_std"] - o[2]) <= 1e-9 * o[2]
