"""GPU NUFFT operators (HIP, gfx950) against the FP64 oracle and the exact DFT.

Tolerances: FP32 operators 2e-5 of the largest output, FP64 operators 1e-9
(the C-24 window is reflected for v<0 points, which moves ~1e-12 of weight).
"""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu


def _case(N, M, seed, amax):
    rng = np.random.default_rng(seed)
    u, v = synth.elliptical_arcs(M, rng, narcs=6, amax=amax)
    x = rng.standard_normal((N, N))
    y = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    return u, v, x, y


@pytest.mark.parametrize("preset", [sa.PRESET_C24, sa.PRESET_CXX12])
@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("N,M", [(64, 1500), (128, 4000), (100, 900)])
def test_forward_adjoint_vs_oracle(preset, fp64, N, M):
    u, v, x, y = _case(N, M, 7 + N, 0.35)
    print(f"case preset={preset} fp64={fp64} N={N} M={M}", flush=True)
    yo = oracle.nufft_forward(u, v, x, preset)
    yg = sa.nufft_forward(u, v, x, preset, fp64=fp64)
    ef = np.abs(yg - yo).max() / np.abs(yo).max()
    xo = oracle.nufft_adjoint(u, v, y, N, N, preset)
    xg = sa.nufft_adjoint(u, v, y, N, N, preset, fp64=fp64)
    ea = np.abs(xg - xo).max() / np.abs(xo).max()
    print(f"  forward rel err {ef:.3e}  adjoint rel err {ea:.3e}")
    tol = 1e-9 if fp64 else 2e-5
    assert ef < tol and ea < tol


def test_adjoint_identity_fp32():
    N, M = 128, 5000
    u, v, x, y = _case(N, M, 3, 0.4)
    Ax = sa.nufft_forward(u, v, x, sa.PRESET_CXX12)
    Ahy = sa.nufft_adjoint(u, v, y, N, N, sa.PRESET_CXX12)
    lhs = np.real(np.vdot(y, Ax))
    rhs = np.sum(x * Ahy)
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_forward_vs_exact_dft():
    N, M = 64, 700
    u, v, x, _ = _case(N, M, 11, 0.35)
    ref = synth.dft_vis(x, u, v)
    for preset, tol in ((sa.PRESET_C24, 1e-9), (sa.PRESET_CXX12, 2e-5)):
        yg = sa.nufft_forward(u, v, x, preset, fp64=True)
        assert np.abs(yg - ref).max() / np.abs(ref).max() < tol


def test_x2y_y2x_exports():
    N, M = 64, 800
    u, v, x, y = _case(N, M, 5, 0.35)
    yr, yi = sa.x2y_nufft(u, v, x)
    yo = oracle.nufft_forward(u, v, x, oracle.PRESET_CXX12)
    assert np.abs((yr + 1j * yi) - yo).max() / np.abs(yo).max() < 1e-9
    xa = sa.y2x_nufft(u, v, y.real, y.imag, N, N)
    xo = oracle.nufft_adjoint(u, v, y, N, N, oracle.PRESET_CXX12)
    assert np.abs(xa - xo).max() / np.abs(xo).max() < 1e-9


def test_fov_violation_is_an_error():
    N = 64
    u = np.array([0.1, 2.9])
    v = np.array([0.1, 0.2])
    with pytest.raises(sa.MfistaError) as e:
        sa.nufft_forward(u, v, np.ones((N, N)))
    assert e.value.status == -3
