"""FP64 CPU restatement of the leave-one-out error estimate -- TEST
INFRASTRUCTURE ONLY.

Follows c/looe_lib.c (Obuchi-Kabashima approximate LOOE for the L1 and
L1+TSV DFT-mode solutions):

- active set S = {i : |x_i| > 0}                          (find_active_set, :52-66)
- A_s = the columns of the column-major M x N matrix A in S   (shrink_A, :70-83)
- H = A_s^T A_s                                           (compute_Hessian_L1, :170-193)
  + lambda_tsv diag(d2_TSV(S))                            (compute_Hessian_L1_TSV, :248-274)
- solve H X = A_s^T (Cholesky, LAPACK dposv "L")          (solve_lin_looe, :87-113)
- d_m = a_m^T H^{-1} a_m, t_m = (y - A x)_m / (1 - d_m), s_m = t_m^2 / 2
- LOOE = mean(s), std = sqrt(sum s^2 / (m-1) - mean^2 m / (m-1))
                                                          (compute_LOOE_core, :115-168)
- H not positive definite: LOOE = -1, looe_m = looe_std = 0.

d2_TSV (:218-235) is the diagonal of the TSV Hessian: 8 in the interior, 6 on
an edge, 4 at a corner, with pixel i at row r = i % NX and column c = i / NX.
The reference tests both r and c against NX (also for the column), and so
does this restatement.
"""
from __future__ import annotations

import numpy as np


def active_set(x):
    """find_active_set (c/looe_lib.c:52-66)."""
    return np.flatnonzero(np.abs(np.asarray(x, dtype=np.float64)) > 0)


def d2_tsv(i, NX, NY):
    """d2_TSV (c/looe_lib.c:218-235); NY is unused there, as here."""
    r, c = i % NX, i // NX
    r_in = 0 < r < NX - 1
    c_in = 0 < c < NX - 1
    if r_in:
        return 8.0 if c_in else 6.0
    return 6.0 if c_in else 4.0


def hessian(A, S, NX=0, NY=0, lambda_tsv=0.0):
    """compute_Hessian_L1 / compute_Hessian_L1_TSV: A_s^T A_s (+ TSV diagonal)."""
    As = np.asarray(A, dtype=np.float64)[:, S]
    H = As.T @ As
    if lambda_tsv > 0:
        H[np.diag_indices_from(H)] += lambda_tsv * np.array([d2_tsv(int(j), NX, NY) for j in S])
    return As, H


def looe_core(As, H, yAx):
    """compute_LOOE_core (c/looe_lib.c:115-168).  Returns (LOOE, looe_m,
    looe_std, leverage d) or (-1, 0, 0, None) if H is not positive definite."""
    yAx = np.asarray(yAx, dtype=np.float64)
    m = yAx.size
    if As.shape[1] > 0:
        try:
            L = np.linalg.cholesky(H)
        except np.linalg.LinAlgError:
            return -1.0, 0.0, 0.0, None
        Y = np.linalg.solve(L, As.T)  # L^{-1} A_s^T, so d_m = |Y[:, m]|^2 = a_m^T H^{-1} a_m
        d = np.sum(Y * Y, axis=0)
    else:
        d = np.zeros(m)
    t = yAx / (1.0 - d)
    s = t * t / 2.0
    looe_m = s.sum() / m
    var = (s * s).sum() / (m - 1.0) - looe_m * looe_m * m / (m - 1.0)
    return looe_m, looe_m, float(np.sqrt(var)), d


def compute_looe(A, x, yAx, NX=0, NY=0, lambda_tsv=0.0):
    """compute_LOOE_L1 (lambda_tsv = 0, c/looe_lib.c:195-216) and
    compute_LOOE_L1_TSV (c/looe_lib.c:276-309).  A: M x N (numpy, any order),
    x: N, yAx = y - A x.  Returns (LOOE, looe_m, looe_std)."""
    S = active_set(x)
    As, H = hessian(A, S, NX, NY, lambda_tsv)
    r = looe_core(As, H, yAx)
    return r[0], r[1], r[2]
