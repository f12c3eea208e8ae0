"""The NUFFT set-up's tile sort (a stable counting sort, sort.hip) on inputs
that stress it: a handful of visibilities (fewer than a wave, exactly a wave,
one more), and a coverage clustered into a few tiles with thousands of
entries each, checked through the operators against the FP64 oracle (the
reference's sort_input, c++/mfista_nufft_lib.cpp:550-641, only orders work;
results must not depend on it).  Bars: forward and adjoint relative error
<= 2e-5 (FP32 operators, 12 taps) and <= 1e-9 (FP64)."""
import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("M", [1, 2, 63, 64, 65, 4097])
def test_operators_small_and_odd_counts(M):
    rng = np.random.default_rng(M)
    N = 64
    u = rng.uniform(-1.2, 1.2, M)
    v = rng.uniform(-1.2, 1.2, M)
    x = rng.standard_normal((N, N))
    r = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    print(f"case sort edges M={M}", flush=True)
    for fp64, tol in ((False, 2e-5), (True, 1e-9)):
        ef = _rel(sa.nufft_forward(u, v, x, preset=sa.PRESET_CXX12, fp64=fp64),
                  oracle.nufft_forward(u, v, x, oracle.PRESET_CXX12))
        ea = _rel(sa.nufft_adjoint(u, v, r, N, N, preset=sa.PRESET_CXX12, fp64=fp64),
                  oracle.nufft_adjoint(u, v, r, N, N, oracle.PRESET_CXX12))
        print(f"  fp64={fp64}: forward {ef:.2e} adjoint {ea:.2e}", flush=True)
        assert ef <= tol and ea <= tol


def test_clustered_coverage():
    """30000 visibilities in a disc of radius 0.02 rad/px: all footprints
    fall into the few tiles around the origin (thousands of entries per tile,
    most tiles empty)."""
    rng = np.random.default_rng(9)
    N, M = 128, 30000
    rad = 0.02 * np.sqrt(rng.uniform(0, 1, M))
    th = rng.uniform(0, 2 * np.pi, M)
    u, v = rad * np.cos(th), rad * np.sin(th)
    x = rng.standard_normal((N, N))
    r = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    print(f"case sort clustered N={N} M={M}", flush=True)
    for preset, op in ((sa.PRESET_CXX12, oracle.PRESET_CXX12), (sa.PRESET_C24, oracle.PRESET_C24)):
        ef = _rel(sa.nufft_forward(u, v, x, preset=preset, fp64=True), oracle.nufft_forward(u, v, x, op))
        ea = _rel(sa.nufft_adjoint(u, v, r, N, N, preset=preset, fp64=True), oracle.nufft_adjoint(u, v, r, N, N, op))
        print(f"  preset {preset}: forward {ef:.2e} adjoint {ea:.2e}", flush=True)
        assert ef <= 1e-9 and ea <= 1e-9
