"""Writes the committed fixtures under tests/golden/.

The reference ships inputs only (no expected outputs), so every expected value
here comes from a closed form computed independently of the oracle:
  * nufft_dft_*.npz   exact FP64 DFT  V_k = sum_ij x_ij exp(+i(u_k(i-N/2)+v_k(j-N/2)))
                      and its adjoint Re(A^H r); u,v,x,r seeded
  * fftmode_16.npz    F(x) = 1/4 sum_full |mask (y - fft2(x)/sqrt(NN))|^2 and -dF/dx
                      from numpy.fft with the reference's half-plane rules
  * reg_5x7.npz       TSV, TV, dTSV of a non-square image in the x[NX*j+i] view
  * ref_nufft_data.npz, ref_fft_data.npz
                      the reference's example data files c/nufft_data.txt and
                      c/fft_data.txt, converted to arrays (read when this script
                      runs in a container that has the reference mounted)
  * ref_data_files.json
                      byte size and SHA-256 of those two text files, so the
                      writers of the data-file formats can be checked to
                      reproduce them byte for byte from the arrays
Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def dft_forward(u, v, x):
    N = x.shape[0]
    p = np.arange(N) - N // 2
    E = np.exp(1j * (u[:, None, None] * p[None, :, None] + v[:, None, None] * p[None, None, :]))
    return np.einsum("kij,ij->k", E, x)


def dft_adjoint(u, v, r, N):
    p = np.arange(N) - N // 2
    E = np.exp(1j * (u[:, None, None] * p[None, :, None] + v[:, None, None] * p[None, None, :]))
    return np.real(np.einsum("kij,k->ij", np.conj(E), r))


def nufft_golden(N, M, seed, umax):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, M)
    v = rng.uniform(-umax, umax, M)
    x = rng.standard_normal((N, N))
    r = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    np.savez_compressed(os.path.join(HERE, f"nufft_dft_{N}.npz"), u=u, v=v, x=x, r=r,
                        Ax=dft_forward(u, v, x), AHr=dft_adjoint(u, v, r, N))


def fftmode_golden():
    rng = np.random.default_rng(16)
    N = 16
    cells = set()
    while len(cells) < 40:
        a, b = rng.integers(-7, 8), rng.integers(0, 8)
        if b > 0 or a > 0:
            cells.add((int(a), int(b)))
    u, v, yr, yi, sd = [], [], [], [], []
    for a, b in sorted(cells):
        val = rng.standard_normal() + 1j * rng.standard_normal()
        s = rng.uniform(0.5, 2.0)
        u += [a % N, (-a) % N]
        v += [b % N, (-b) % N]
        yr += [val.real, val.real]
        yi += [val.imag, -val.imag]
        sd += [s, s]
    u, v = np.array(u, np.int32), np.array(v, np.int32)
    yr, yi, sd = np.array(yr), np.array(yi), np.array(sd)
    x = rng.standard_normal((N, N))
    # idx2mat (later entries overwrite), then the full-grid residual of the half plane
    mask = np.zeros((N, N))
    yf = np.zeros((N, N), complex)
    for k in range(u.size):
        mask[u[k], v[k]] = 1 / sd[k]
        yf[u[k], v[k]] = (yr[k] + 1j * yi[k]) / sd[k]
    X = np.fft.fft2(x) / N
    h = N // 2 + 1
    rh = yf[:, :h] - mask[:, :h] * X[:, :h]
    w = np.full(h, 2.0)
    w[0] = w[-1] = 1.0
    F = 0.25 * np.sum(w[None, :] * np.abs(rh) ** 2)
    g = np.fft.irfft2(mask[:, :h] * rh / (2 * N), s=(N, N)) * N * N
    np.savez_compressed(os.path.join(HERE, "fftmode_16.npz"), u=u, v=v, yr=yr, yi=yi, sd=sd, x=x,
                        F=F, g=g)


def reg_golden():
    rng = np.random.default_rng(57)
    NX, NY = 5, 7
    x = rng.standard_normal(NX * NY)
    im = x.reshape(NY, NX)  # row j, column i  <->  x[NX*j+i]
    dh = im[:, :-1] - im[:, 1:]
    dv = im[:-1, :] - im[1:, :]
    tsv = np.sum(dh ** 2) + np.sum(dv ** 2)
    tv = np.sum(np.sqrt(dh[:-1, :] ** 2 + dv[:, :-1] ** 2)) + np.sum(np.abs(dh[-1, :])) + np.sum(np.abs(dv[:, -1]))
    g = np.zeros_like(im)
    g[:, :-1] += 2 * dh
    g[:, 1:] -= 2 * dh
    g[:-1, :] += 2 * dv
    g[1:, :] -= 2 * dv
    np.savez_compressed(os.path.join(HERE, "reg_5x7.npz"), x=x, NX=NX, NY=NY, tsv=tsv, tv=tv,
                        dtsv=g.ravel())


def _read_table(path, kinds):
    with open(path) as f:
        lines = f.read().splitlines()
    hdr = {}
    for ln in lines[:3]:
        k, v = ln.split("=")
        hdr[k.strip()] = int(v)
    rows = [ln for ln in lines[4:] if ln.strip() and "," in ln and not ln.strip()[0].isalpha()]
    cols = list(zip(*[r.split(",") for r in rows]))
    return hdr, [np.array([k(c) for c in col]) for k, col in zip(kinds, cols)]


def reference_data():
    p = os.path.join(REF, "c", "nufft_data.txt")
    if os.path.exists(p):
        hdr, (u, v, vr, vi, sd) = _read_table(p, [float] * 5)
        assert u.size == hdr["M"]
        np.savez_compressed(os.path.join(HERE, "ref_nufft_data.npz"), u=u, v=v, vis_r=vr, vis_i=vi,
                            std=sd, NX=hdr["NX"], NY=hdr["NY"])
    p = os.path.join(REF, "c", "fft_data.txt")
    if os.path.exists(p):
        hdr, (u, v, yr, yi, sd) = _read_table(p, [int, int, float, float, float])
        assert u.size == hdr["M"]
        np.savez_compressed(os.path.join(HERE, "ref_fft_data.npz"), u=u.astype(np.int32),
                            v=v.astype(np.int32), y_r=yr, y_i=yi, noise=sd, NX=hdr["NX"], NY=hdr["NY"])


def reference_file_digests():
    import hashlib
    import json

    out = {}
    for name in ("nufft_data.txt", "fft_data.txt"):
        p = os.path.join(REF, "c", name)
        if os.path.exists(p):
            b = open(p, "rb").read()
            out[name] = {"bytes": len(b), "sha256": hashlib.sha256(b).hexdigest()}
    if out:
        with open(os.path.join(HERE, "ref_data_files.json"), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
            f.write("\n")


def oracle_refdata_solutions():
    """Optimum of the reference's nufft_data.txt problem with the author's run
    parameters (matlab/testdata4nufft.mat: lambda_l1 = 100, lambda_tsv = 100,
    nonneg), for both NUFFT presets.

    With nonneg the L1 term is linear (lambda_l1 sum x), so the objective
    chi^2/2 + lambda_l1 sum x + lambda_tsv TSV(x) is smooth on x >= 0 and
    scipy's L-BFGS-B finds its optimum independently of MFISTA, with the
    oracle's FP64 forward and adjoint NUFFT (pinned by the exact-DFT vectors)
    as function and gradient.  MFISTA's own stopping rule (relative change
    over 50 iterations) stops early on this flat objective (2e-6 .. 8e-6
    above the optimum), so it is not used here.  Starts from the oracle's
    MFISTA solution at a 1e-8 tolerance; 6000 L-BFGS-B iterations (minutes
    per preset)."""
    import sys

    from scipy.optimize import minimize

    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import oracle

    d = np.load(os.path.join(HERE, "ref_nufft_data.npz"))
    N = int(d["NX"])
    u, v = d["u"], d["v"]
    V = d["vis_r"] + 1j * d["vis_i"]
    w2 = 1.0 / d["std"] ** 2
    l1 = ltsv = 100.0
    out = {}
    for preset in (oracle.PRESET_CXX12, oracle.PRESET_C24):
        o = oracle.imaging_core_nufft(u, v, d["vis_r"], d["vis_i"], d["std"], N, N, 20000, 0.0, l1, 0.0, ltsv,
                                      5e8, nonneg=1, preset=preset, rel_tol=1e-8)

        def fg(x):
            r = V - oracle.nufft_forward(u, v, x.reshape(N, N), preset=preset)
            f = 0.5 * np.sum(w2 * np.abs(r) ** 2) + l1 * x.sum() + ltsv * oracle.tsv_flat(x, N, N)
            g = (-oracle.nufft_adjoint(u, v, w2 * r, N, N, preset=preset).ravel() + l1
                 + ltsv * oracle.d_tsv_flat(x, N, N))
            return f, g

        r = minimize(fg, o["x"].ravel(), jac=True, method="L-BFGS-B", bounds=[(0, None)] * (N * N),
                     options={"maxiter": 6000, "maxfun": 12000, "ftol": 1e-17, "gtol": 1e-9, "maxcor": 30})
        x = r.x.reshape(N, N)
        res = oracle.result_nufft(u, v, d["vis_r"], d["vis_i"], d["std"], x, l1, 0.0, ltsv, preset=preset)
        out[f"x{preset}"] = x.astype(np.float32)
        out[f"finalcost{preset}"] = res["finalcost"]
        out[f"mfista_finalcost{preset}"] = o["result"]["finalcost"]
        print("preset", preset, "L-BFGS-B", res["finalcost"], "MFISTA(1e-8)", o["result"]["finalcost"], r.nit,
              flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_nufft_solution.npz"), **out)


if __name__ == "__main__":
    import sys as _sys

    if len(_sys.argv) > 1 and _sys.argv[1] == "solutions":
        oracle_refdata_solutions()
        raise SystemExit(0)
    nufft_golden(32, 40, 3201, 1.5)
    nufft_golden(16, 25, 1601, 0.5)
    fftmode_golden()
    reg_golden()
    reference_data()
    reference_file_digests()
    print("fixtures written to", HERE)
