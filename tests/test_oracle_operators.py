"""Pins the FP64 oracle's operators to closed forms (CPU only).

* FFT: numpy.fft (FFTW's sign and c2r conventions)
* NUFFT forward/adjoint: the exact DFT, committed as golden vectors
  (tests/golden/nufft_dft_*.npz, written by make_golden.py); C-24 must match to
  1e-9, CXX-12 (12 taps) to 2e-5 -- a dropped E3 or E4 factor, a sign flip
  or a transposed index fails these by orders of magnitude
* invariants: adjoint identity, Hermitian symmetry, linearity
"""
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("shape", [(8, 8), (16, 12), (10, 15), (400, 400), (30, 20)])
def test_r2c_matches_numpy(shape):
    x = np.random.default_rng(1).standard_normal(shape)
    h = oracle.fft2_r2c(x)
    ref = np.fft.rfft2(x)
    assert np.abs(h - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(8, 8), (12, 10), (9, 14), (50, 40)])
def test_c2r_matches_numpy_irfft(shape):
    """c2r of an arbitrary (non-Hermitian) half spectrum: numpy's irfft2 also
    transforms the leading axis first and drops Im of the DC/Nyquist columns."""
    rng = np.random.default_rng(2)
    n0, n1 = shape
    h = rng.standard_normal((n0, n1 // 2 + 1)) + 1j * rng.standard_normal((n0, n1 // 2 + 1))
    out = oracle.fft2_c2r(h, n1)
    ref = np.fft.irfft2(h, s=shape) * n0 * n1
    assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 97, 400, 1000])
def test_c2c_both_signs(n):
    z = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    assert np.allclose(oracle.fft_c2c(z, -1), np.fft.fft(z), atol=1e-10 * max(1, n))
    assert np.allclose(oracle.fft_c2c(z, +1), np.fft.ifft(z) * n, atol=1e-10 * max(1, n))


@pytest.mark.parametrize("N", [16, 32])
@pytest.mark.parametrize("preset,tol", [(oracle.PRESET_C24, 1e-9), (oracle.PRESET_CXX12, 2e-5)])
def test_nufft_vs_golden_dft(N, preset, tol):
    d = np.load(os.path.join(G, f"nufft_dft_{N}.npz"))
    y = oracle.nufft_forward(d["u"], d["v"], d["x"], preset)
    assert np.abs(y - d["Ax"]).max() / np.abs(d["Ax"]).max() < tol
    a = oracle.nufft_adjoint(d["u"], d["v"], d["r"], N, N, preset)
    assert np.abs(a - d["AHr"]).max() / np.abs(d["AHr"]).max() < tol


def test_nufft_golden_is_not_trivial():
    """The tolerance above must separate right from wrong: the conjugate of the
    DFT (a sign mistake) misses the golden vector by O(1)."""
    d = np.load(os.path.join(G, "nufft_dft_32.npz"))
    y = oracle.nufft_forward(d["u"], d["v"], d["x"], oracle.PRESET_C24)
    assert np.abs(np.conj(y) - d["Ax"]).max() / np.abs(d["Ax"]).max() > 0.1


@pytest.mark.parametrize("preset", [oracle.PRESET_C24, oracle.PRESET_CXX12])
def test_nufft_invariants(preset):
    rng = np.random.default_rng(9)
    N, M = 48, 300
    u = rng.uniform(-1.5, 1.5, M)
    v = rng.uniform(-1.5, 1.5, M)
    x = rng.standard_normal((N, N))
    x2 = rng.standard_normal((N, N))
    r = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    Ax = oracle.nufft_forward(u, v, x, preset)
    # adjoint identity <A x, r>_R = <x, Re A^H r>
    lhs = np.real(np.vdot(r, Ax))
    rhs = np.sum(x * oracle.nufft_adjoint(u, v, r, N, N, preset))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # V(-u,-v) = conj V(u,v) for a real image (up to the tap approximation)
    Am = oracle.nufft_forward(-u, -v, x, preset)
    tol = 1e-9 if preset == oracle.PRESET_C24 else 2e-5
    assert np.abs(Am - np.conj(Ax)).max() <= tol * np.abs(Ax).max()
    # linearity
    A2 = oracle.nufft_forward(u, v, 2 * x - 3 * x2, preset)
    A1 = oracle.nufft_forward(u, v, x2, preset)
    assert np.abs(A2 - (2 * Ax - 3 * A1)).max() <= 1e-12 * np.abs(A2).max() * 10


def test_fov_violation_rejected():
    with pytest.raises(ValueError):
        oracle.nufft_forward(np.array([3.0]), np.array([0.0]), np.ones((32, 32)))


def test_fftmode_vs_golden():
    d = np.load(os.path.join(G, "fftmode_16.npz"))
    F, g = oracle.fft_F_grad(d["u"], d["v"], d["yr"], d["yi"], d["sd"], d["x"])
    assert abs(F - float(d["F"])) <= 1e-12 * abs(float(d["F"]))
    assert np.abs(g - d["g"]).max() <= 1e-12 * np.abs(d["g"]).max()


def test_fftmode_gradient_is_minus_dF():
    """-g is the gradient of F (finite differences on the closed-form F)."""
    d = np.load(os.path.join(G, "fftmode_16.npz"))
    x = d["x"].copy()
    F0, g = oracle.fft_F_grad(d["u"], d["v"], d["yr"], d["yi"], d["sd"], x)
    rng = np.random.default_rng(4)
    for _ in range(5):
        i, j = rng.integers(0, 16, 2)
        h = 1e-6
        xp = x.copy()
        xp[i, j] += h
        xm = x.copy()
        xm[i, j] -= h
        Fp, _ = oracle.fft_F_grad(d["u"], d["v"], d["yr"], d["yi"], d["sd"], xp)
        Fm, _ = oracle.fft_F_grad(d["u"], d["v"], d["yr"], d["yi"], d["sd"], xm)
        assert abs((Fp - Fm) / (2 * h) + g[i, j]) <= 1e-6 * max(1.0, abs(g[i, j]))
