"""The LOOE oracle (oracle/looe_ref.py, a restatement of c/looe_lib.c) pinned
on the CPU:

- the hat-matrix identity: with every pixel active and lambda_tsv = 0,
  (y - A x)_m / (1 - d_m) is exactly the residual of the least-squares fit
  that leaves observation m out (brute force: refit without row m);
- the TSV Hessian diagonal against the Hessian of the TSV function itself
  (finite second differences), and the reference's column-against-NX test;
- the mean and standard deviation against numpy's sample statistics;
- a Hessian that is not positive definite returns -1 (looe_m = looe_std = 0),
  as compute_LOOE_core does when dposv fails."""
import numpy as np

from oracle import looe_ref as lr


def _tsv(x, NX, NY):
    img = x.reshape(NY, NX)  # pixel i at row i % NX, column i // NX
    return float(np.sum((img[:, :-1] - img[:, 1:]) ** 2) + np.sum((img[:-1, :] - img[1:, :]) ** 2))


def test_hat_matrix_identity_is_exact_leave_one_out():
    rng = np.random.default_rng(3)
    M, N = 40, 12
    A = rng.standard_normal((M, N))
    y = A @ rng.standard_normal(N) + 0.1 * rng.standard_normal(M)
    x = np.linalg.lstsq(A, y, rcond=None)[0]  # every pixel active
    yAx = y - A @ x
    _, _, _, d = lr.looe_core(*lr.hessian(A, lr.active_set(x)), yAx)
    t = yAx / (1 - d)
    for m in range(M):
        keep = np.arange(M) != m
        xm = np.linalg.lstsq(A[keep], y[keep], rcond=None)[0]
        assert abs((y[m] - A[m] @ xm) - t[m]) <= 1e-10 * max(1.0, abs(t[m]))
    # LOOE = mean of t^2 / 2 over the observations
    looe, m_, s_ = lr.compute_looe(A, x, yAx)
    assert abs(looe - np.mean(t * t / 2)) <= 1e-12 * looe
    assert abs(s_ - np.std(t * t / 2, ddof=1)) <= 1e-9 * s_


def test_tsv_hessian_diagonal():
    NX = NY = 5
    x0 = np.random.default_rng(1).standard_normal(NX * NY)
    h = 1e-3
    for i in range(NX * NY):
        e = np.zeros(NX * NY)
        e[i] = h
        fd = (_tsv(x0 + e, NX, NY) - 2 * _tsv(x0, NX, NY) + _tsv(x0 - e, NX, NY)) / (h * h)
        assert abs(fd - lr.d2_tsv(i, NX, NY)) <= 1e-6, (i, fd)  # d^2 TSV / dx_i^2
    # corners 4, edges 6, interior 8 (the second derivative of TSV itself)
    assert [lr.d2_tsv(i, 5, 5) for i in (0, 4, 20, 24)] == [4.0] * 4
    assert lr.d2_tsv(2, 5, 5) == 6.0 and lr.d2_tsv(10, 5, 5) == 6.0 and lr.d2_tsv(12, 5, 5) == 8.0
    # the reference tests the column index against NX: on a 4 x 6 image
    # (NX = 4) column 3 already counts as the last one
    assert lr.d2_tsv(1 + 4 * 3, 4, 6) == 6.0 and lr.d2_tsv(1 + 4 * 2, 4, 6) == 8.0


def test_tsv_hessian_added_to_active_diagonal():
    rng = np.random.default_rng(5)
    NX, NY, M = 4, 4, 30
    A = rng.standard_normal((M, NX * NY))
    x = np.zeros(NX * NY)
    x[[0, 5, 6, 15]] = [1.0, 2.0, 0.5, 3.0]
    S = lr.active_set(x)
    assert list(S) == [0, 5, 6, 15]
    As, H = lr.hessian(A, S, NX, NY, 2.5)
    ref = A[:, S].T @ A[:, S] + np.diag([2.5 * 4, 2.5 * 8, 2.5 * 8, 2.5 * 4])
    assert np.allclose(H, ref, rtol=0, atol=1e-12)
    # the leverage is a_m^T H^{-1} a_m
    _, _, _, d = lr.looe_core(As, H, np.ones(M))
    assert np.allclose(d, np.einsum("mi,ij,mj->m", As, np.linalg.inv(H), As), rtol=1e-10)


def test_not_positive_definite_and_empty_active_set():
    rng = np.random.default_rng(7)
    M, N = 5, 8  # more active pixels than observations: A_s^T A_s is singular
    A = rng.standard_normal((M, N))
    x = np.ones(N)
    looe, m_, s_ = lr.compute_looe(A, x, rng.standard_normal(M))
    assert (looe, m_, s_) == (-1.0, 0.0, 0.0)
    # no active pixel: d = 0, LOOE = mean((y - A x)^2 / 2)
    r = rng.standard_normal(M)
    looe, m_, s_ = lr.compute_looe(A, np.zeros(N), r)
    assert abs(looe - np.mean(r * r / 2)) <= 1e-14
