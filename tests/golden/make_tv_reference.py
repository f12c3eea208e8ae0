"""Writes tests/golden/tv_oracle.npz: the FP64 oracle's L1+TV solutions (24-tap
preset, FGP 50 iterations; c/mfista_nufft_lib.c:308-553 with the FGP prox of
c/mfista_TV_lib.c restated in oracle/mfista_oracle.c) after a fixed 20000
MFISTA iterations, for the two GPU test cases of
tests/test_gpu_signed_entry.py::test_nufft_tv_signed_and_box:
  fgp_l1  nonneg_flag = 0 (FGP_L1) on a sky with negative regions
  box     nonneg_flag = 1 with a CLEAN box (FGP_nonneg_box)
Both 64^2, M = 2500, lambda_l1 = 0.5, lambda_tv = 2, cinit = 1e4.  The
signed TV problem converges slowly, so the comparison is made far past the
point where a stopping rule would end it.  Stored: image, objective and the
SHA-256 of the inputs per case.  Takes about 15 minutes on one core per case.
Run:  python tests/golden/make_tv_reference.py [case]
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from sparseimaging_amd import synth  # noqa: E402

ITERS = 20000
KW = dict(lambda_l1=0.5, lambda_tv=2.0, cinit=1e4, fgp_iter=50)


def signed_problem(N, M, seed):
    """A sky with negative regions, 12 arcs (the same generator as the GPU test)."""
    rng = np.random.default_rng(seed)
    x = synth.gaussian_blobs(N, rng, 3) - 0.7 * synth.gaussian_blobs(N, rng, 2)
    u, v = synth.elliptical_arcs(M, rng, amax=0.35)
    vis = synth.dft_vis(x, u, v)
    sig = 0.01 * np.abs(vis).max()
    vis = vis + sig * (rng.standard_normal(M) + 1j * rng.standard_normal(M))
    return {"u": u, "v": v, "vis_r": vis.real.copy(), "vis_i": vis.imag.copy(), "std": np.full(M, sig)}


def case_inputs(case):
    if case == "fgp_l1":
        return signed_problem(64, 2500, 72), 0, None
    P = synth.make_nufft_problem(N=64, M=2500, seed=73, amax=0.35, kind="tv", coverage="disc")
    box = np.ones((64, 64), np.float32)
    box[:, :12] = 0.0
    box[50:, :] = 0.0
    return P, 1, box


def sha(P):
    h = hashlib.sha256()
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        h.update(np.ascontiguousarray(P[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def run(case):
    P, nonneg, box = case_inputs(case)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, ITERS, 0.0,
                                  KW["lambda_l1"], KW["lambda_tv"], 0.0, KW["cinit"], nonneg=nonneg,
                                  cl_box=None if box is None else box.ravel(), preset=oracle.PRESET_C24,
                                  fgp_iter=KW["fgp_iter"], rel_tol=0.0)
    print(f"{case}: ITER {o['result']['ITER']} finalcost {o['result']['finalcost']!r}", flush=True)
    return {f"{case}_x": o["x"], f"{case}_F": np.array(o["result"]["finalcost"]), f"{case}_sha256": np.array(sha(P))}


def main():
    cases = sys.argv[1:] or ["fgp_l1", "box"]
    out = {}
    path = os.path.join(HERE, "tv_oracle.npz")
    if os.path.exists(path):
        out.update(dict(np.load(path)))
    for c in cases:
        out.update(run(c))
    np.savez_compressed(path, **out)


if __name__ == "__main__":
    main()
