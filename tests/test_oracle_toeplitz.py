"""The Toeplitz identity the Toeplitz loop applies, checked with the FP64
oracle on the CPU (no GPU work).

For a real image x, Re A^H W^2 A x is a convolution with the point-spread
function p(D) = sum_m w_m^2 cos(u_m . D); embedded in the circulant of the
2N x 2N grid its spectrum is real, and b = Re A^H W^2 V is the adjoint of
w^2 V.  The loop takes p and b from the 24-tap adjoint on a 2N x 2N image
(toeplitz.hip); here the same construction in numpy, from the oracle's
adjoint, is compared with the oracle's own A^H W^2 A x, for both presets, and
with the exact DFT.  A wrong centring, a transposed offset or a dropped weight
fails these by orders of magnitude.  The last tests check the even/odd split
the loop's pruned 2N-point transforms rest on against the plain padded
convolution (numpy FFTs only).
"""
import numpy as np
import pytest

import oracle as O


def _problem(N, M, seed):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0.0025, 0.3, M)) * np.pi * 0.8  # inside the 24-tap field of view at N = 32
    th = rng.uniform(0, 2 * np.pi, M)
    u, v = r * np.cos(th), r * np.sin(th)
    w = 1.0 / rng.uniform(0.5, 2.0, M)
    V = rng.normal(size=M) + 1j * rng.normal(size=M)
    x = np.abs(rng.normal(size=(N, N)))
    return u, v, w, V, x


def _toeplitz_apply(p2, x):
    """T x from the 2N x 2N point-spread image p2 (pixel (I, J) at offset
    (I - N, J - N)), as the loop applies it: circulant embedding with the
    offset -N row and column set to 0, a real spectrum, an FFT convolution of
    the zero-padded image, the N x N crop."""
    N = x.shape[0]
    L = 2 * N
    c = np.zeros((L, L))
    for a in range(L):
        if a == N:
            continue
        dx = a if a < N else a - L
        for b in range(L):
            if b == N:
                continue
            dy = b if b < N else b - L
            c[a, b] = p2[dx + N, dy + N]
    K = np.fft.fft2(c)
    assert np.abs(K.imag).max() <= 1e-9 * np.abs(K.real).max()  # the embedding is even
    xp = np.zeros((L, L))
    xp[:N, :N] = x
    return np.real(np.fft.ifft2(np.fft.fft2(xp) * K.real))[:N, :N]


@pytest.mark.parametrize("preset,bar", [(O.PRESET_C24, 1e-9), (O.PRESET_CXX12, 1e-4)])
def test_toeplitz_matches_normal_operator(preset, bar):
    N, M = 32, 800
    u, v, w, V, x = _problem(N, M, 11)
    y = O.nufft_forward(u, v, x, preset=preset)
    T_ref = O.nufft_adjoint(u, v, w ** 2 * y, N, N, preset=preset)
    p2 = O.nufft_adjoint(u, v, (w ** 2).astype(complex), 2 * N, 2 * N, preset=O.PRESET_C24)
    Tx = _toeplitz_apply(p2, x)
    err = np.linalg.norm(Tx - T_ref) / np.linalg.norm(T_ref)
    # 24 taps: the same operator to 1e-11; 12 taps: its own approximation error (~1e-6)
    assert err <= bar


def test_toeplitz_matches_exact_dft_and_b():
    N, M = 32, 800
    u, v, w, V, x = _problem(N, M, 12)
    i = np.arange(N) - N // 2
    E = np.exp(1j * (np.outer(u, np.repeat(i, N)) + np.outer(v, np.tile(i, N))))  # forward: e^{+i(u i + v j)}
    T_exact = np.real(E.conj().T @ (w ** 2 * (E @ x.ravel()))).reshape(N, N)
    p2 = O.nufft_adjoint(u, v, (w ** 2).astype(complex), 2 * N, 2 * N, preset=O.PRESET_C24)
    assert np.linalg.norm(_toeplitz_apply(p2, x) - T_exact) / np.linalg.norm(T_exact) <= 1e-9
    # b = Re A^H W^2 V: the N x N crop of the 2N image's adjoint (pixel (i, j) at offset (i - N/2, j - N/2))
    b_ref = np.real(E.conj().T @ (w ** 2 * V)).reshape(N, N)
    b2 = O.nufft_adjoint(u, v, w ** 2 * V, 2 * N, 2 * N, preset=O.PRESET_C24)
    b = b2[N // 2:N // 2 + N, N // 2:N // 2 + N]
    assert np.linalg.norm(b - b_ref) / np.linalg.norm(b_ref) <= 1e-9
    # chi^2 from T and b: |W V|^2 - 2 <x, b> + <x, T x> = |W (V - A x)|^2
    chi_direct = np.sum(np.abs(w * (V - E @ x.ravel())) ** 2)
    chi_toep = np.sum(np.abs(w * V) ** 2) - 2 * np.sum(x * b_ref) + np.sum(x * T_exact)
    assert abs(chi_toep - chi_direct) <= 1e-10 * chi_direct


def test_rectangular_preset_is_not_the_exact_transform():
    """Why the Toeplitz loop is limited to square images: for NX != NY both
    presets size the second axis's kernel from NX, and the oracle's forward
    transform is then far from the exact DFT."""
    NX, NY, M = 32, 48, 600
    rng = np.random.default_rng(13)
    r = np.sqrt(rng.uniform(0.0025, 0.5, M)) * np.pi * 0.7
    th = rng.uniform(0, 2 * np.pi, M)
    u, v = r * np.cos(th), r * np.sin(th)
    x = np.abs(rng.normal(size=(NX, NY)))
    ii, jj = np.meshgrid(np.arange(NX) - NX // 2, np.arange(NY) - NY // 2, indexing="ij")
    ye = np.exp(1j * (np.outer(u, ii.ravel()) + np.outer(v, jj.ravel()))) @ x.ravel()
    for preset in (O.PRESET_CXX12, O.PRESET_C24):
        y = O.nufft_forward(u, v, x, preset=preset)
        assert np.linalg.norm(y - ye) / np.linalg.norm(ye) > 1e-2


def _pruned_conv(x, K):
    """The loop's pruned 2N-point convolution of one padded row or column
    (st_colconv1, st_rowfft_fwd_h / st_rowfft_c2r_h): pixel i at
    a = (i - N//2) mod 2N, one per residue class r = a mod N; the even and
    odd outputs of the 2N-point transform are the N-point transforms of y(r)
    and y(r) W_2N^a; the N wanted outputs of the backward transform are
    e(r) + conj(W_2N^a) o(r).  Written from those formulas, not from the
    kernels' code."""
    N = x.size
    h = N // 2
    L = 2 * N
    tw = np.exp(-2j * np.pi * np.arange(L) / L)
    r = np.arange(N)
    i = (r + h) % N
    a = np.where(r < h, r, r + N)
    y = x[i]
    X = np.empty(L, complex)
    X[0::2] = np.fft.fft(y)
    X[1::2] = np.fft.fft(y * tw[a])
    P = X * K
    e, o = np.fft.ifft(P[0::2]) * N, np.fft.ifft(P[1::2]) * N
    out = np.empty(N, complex)
    out[i] = e + o * np.conj(tw[a])
    return out


def _padded_conv(x, K):
    N = x.size
    L = 2 * N
    z = np.zeros(L, complex)
    z[(np.arange(N) - N // 2) % L] = x
    full = np.fft.ifft(np.fft.fft(z) * K) * L
    return full[(np.arange(N) - N // 2) % L]


@pytest.mark.parametrize("N", [16, 48, 96, 256, 320, 384, 512])
def test_pruned_transforms_equal_the_padded_convolution(N):
    """The identity behind the loop's pruned transforms holds for the even
    sizes the loop takes (powers of two and the radix-3 / radix-5 sizes),
    against the direct 2N-point convolution of the zero-padded sequence."""
    rng = np.random.default_rng(N)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    K = rng.normal(size=2 * N)
    K = 0.5 * (K + K[(-np.arange(2 * N)) % (2 * N)])  # real and even, as the loop's Kh
    ref = _padded_conv(x, K)
    assert np.abs(_pruned_conv(x, K) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pruned_placement_needs_even_sizes():
    """For odd N the placement a = r for r < N//2 misses the residue class
    r = N//2 (pixel N-1 sits at a = N//2, not at N//2 + N), so the identity
    fails: the loop takes even sizes only."""
    N = 45
    rng = np.random.default_rng(3)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    K = rng.normal(size=2 * N)
    K = 0.5 * (K + K[(-np.arange(2 * N)) % (2 * N)])
    ref = _padded_conv(x, K)
    assert np.abs(_pruned_conv(x, K) - ref).max() > 1e-3 * np.abs(ref).max()


def test_step_quadratic_from_the_half_spectrum():
    """The identities a two-barrier round would use instead of the c2r of both
    trials: <d, T d> = sum_k Kh(k) |D(k)|^2 over the padded grid (D the
    transform of the padded step, Kh the circulant's spectrum over (2N)^2),
    taken over the half spectrum ky = 0 .. N with weights 1, 2, ..., 2, 1;
    and chi^2(z + d) - chi^2(z) = 2 <d, T z - b> + <d, T d>.  T is the
    embedding of p(D) = sum_m w_m^2 cos(u_m . D) from the exact sums, checked
    first against the exact normal operator Re A^H W^2 A of the direct DFT."""
    N, M = 16, 300
    u, v, w, V, _ = _problem(N, M, 21)
    rng = np.random.default_rng(5)
    z = np.abs(rng.normal(size=(N, N)))
    d = rng.normal(size=(N, N))
    L = 2 * N
    D = np.arange(-N, N)
    p2 = np.einsum("m,ma,mb->ab", w ** 2, np.exp(1j * np.outer(u, D)), np.exp(1j * np.outer(v, D))).real
    T = lambda x: _toeplitz_apply(p2, x)  # noqa: E731
    E, F = np.exp(1j * np.outer(u, np.arange(N))), np.exp(1j * np.outer(v, np.arange(N)))
    y = np.einsum("ij,mi,mj->m", z, E.conj(), F.conj())
    Tz_direct = np.einsum("m,mi,mj->ij", w ** 2 * y, E, F).real
    assert np.linalg.norm(T(z) - Tz_direct) <= 1e-12 * np.linalg.norm(Tz_direct)
    c = np.zeros((L, L))
    for a in range(L):
        for b in range(L):
            if a != N and b != N:
                c[a, b] = p2[(a if a < N else a - L) + N, (b if b < N else b - L) + N]
    Kh = np.fft.fft2(c).real / (L * L)
    dp = np.zeros((L, L))
    dp[:N, :N] = d
    Dk = np.fft.fft2(dp)
    full = (Kh * np.abs(Dk) ** 2).sum()
    wgt = np.full(L, 2.0)
    wgt[0] = wgt[N] = 1.0
    wgt[N + 1:] = 0.0  # the half spectrum along the row transform's axis
    half = (Kh * np.abs(Dk) ** 2 * wgt[None, :]).sum()
    dTd = (d * T(d)).sum()
    assert abs(full - dTd) <= 1e-10 * abs(dTd)
    assert abs(half - dTd) <= 1e-10 * abs(dTd)
    # chi^2(x) - |W V|^2 = <x, T x> - 2 <x, b>
    b = rng.normal(size=(N, N))
    q = lambda x: (x * T(x)).sum() - 2 * (x * b).sum()  # noqa: E731
    lhs = q(z + d) - q(z)
    rhs = 2 * (d * (T(z) - b)).sum() + dTd
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)
