"""The command-line tools end to end on the GPU: data file in, image file and
report out, against the library entry points on the same inputs (which the
other GPU tests hold against the oracle) and, for the NUFFT tool, against
the oracle directly."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import sparseimaging_amd as sa
from sparseimaging_amd import datafiles as df
from sparseimaging_amd import synth

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")


def _run(tool, args):
    print(f"case {tool} {' '.join(os.path.basename(str(a)) for a in args)}", flush=True)
    r = subprocess.run([df.tool_path(tool), *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _report_value(text, label):
    for ln in text.splitlines():
        if ln.startswith(" " + label):
            return ln.split(":", 1)[1].strip()
    raise AssertionError(label)


def test_nufft_tool_against_library_and_oracle(tmp_path):
    P = synth.make_nufft_problem(N=64, M=1500, seed=8, amax=0.35)
    data = tmp_path / "vis.txt"
    df.write_nufft_data(data, P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64)
    box = np.ones((64, 64), np.float32)
    box[:, :6] = 0
    (tmp_path / "box.bin").write_bytes(box.tobytes())
    xo, log = tmp_path / "x.out", tmp_path / "run.log"
    # run to a tight relative tolerance: the tool's process may get the stream
    # engine (a second process cannot always take the cooperative launch), whose
    # trajectory differs from the persistent engine's; converged, the two engines
    # agree to the parity bar (test_gpu_solve_nufft.py::test_engines_agree)
    out = _run("mfista_imaging_nufft", [data, 1.0, 0, 10.0, 1e4, xo, "-nonneg", "-maxiter", 20000, "-cl_box",
                                        tmp_path / "box.bin", "-preset", "cxx12", "-rel_tol", 1e-10, "-log", log])
    x = df.read_vector(xo, 64 * 64).reshape(64, 64)
    kw = dict(maxiter=20000, eps=0.0, lambda_l1=1.0, lambda_tsv=10.0, cinit=1e4)
    xl, rl, _ = sa.mfista_imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64,
                                             nonneg_flag=1, cl_box=box, preset=sa.PRESET_CXX12, rel_tol=1e-10, **kw)
    o = oracle.imaging_core_nufft(P["u"], P["v"], P["vis_r"], P["vis_i"], P["std"], 64, 64, kw["maxiter"], kw["eps"],
                                  kw["lambda_l1"], 0.0, kw["lambda_tsv"], kw["cinit"], nonneg=1,
                                  preset=oracle.PRESET_CXX12, cl_box=box, rel_tol=1e-10)
    e_lib = np.linalg.norm(x - xl) / np.linalg.norm(xl)
    e_or = np.linalg.norm(x - o["x"]) / np.linalg.norm(o["x"])
    print(f"  tool vs library relL2 {e_lib:.2e}, vs oracle {e_or:.2e}", flush=True)
    assert e_lib <= 1e-3 and e_or <= 1e-3
    assert np.all(x[:, :6] == 0)
    rep = log.read_text()
    assert rep in out  # the log holds the same two report blocks as stdout
    assert abs(float(_report_value(rep, "cost:")) - rl["finalcost"]) <= 1e-6 * rl["finalcost"]
    assert int(_report_value(rep, "size of input vector:")) == 1500
    assert _report_value(rep, "size of image:") == "64 x 64"


def test_fft_tool_reference_file_tv(tmp_path):
    d = np.load(os.path.join(G, "ref_fft_data.npz"))
    N = int(d["NX"])
    data = tmp_path / "fft_data.txt"
    df.write_fft_data(data, d["u"], d["v"], d["y_r"], d["y_i"], d["noise"], N, N, digits=6)
    xo = tmp_path / "x.out"
    out = _run("mfista_imaging_fft", [data, 10.0, 50.0, 0, 1e4, xo, "-nonneg", "-maxiter", 60, "-fgp", 30,
                                      "-fftw_measure"])
    x = df.read_vector(xo, N * N).reshape(N, N)
    xl, rl, _ = sa.mfista_imaging_core_fft(d["u"], d["v"], d["y_r"], d["y_i"], d["noise"], N, N, maxiter=60,
                                           eps=1e-5, lambda_l1=10.0, lambda_tv=50.0, lambda_tsv=0.0, cinit=1e4,
                                           nonneg_flag=1, fgp_iter=30)
    e = np.linalg.norm(x - xl) / np.linalg.norm(xl)
    print(f"  fft tool vs library relL2 {e:.2e}", flush=True)
    assert e <= 1e-12
    assert int(_report_value(out, "# of iterations:")) == rl["ITER"]
    assert "Lambda_TV:" in out


def test_dft_tool_row_major_rectangular(tmp_path):
    """-t (row-major A) and -rec NX (a 12 x 20 image): the tool must solve the
    same problem as the library given the column-major A."""
    rng = np.random.default_rng(11)
    NX, NY, M = 12, 20, 300
    N = NX * NY
    A = rng.standard_normal((M, N)) / np.sqrt(M)
    y = A @ np.maximum(rng.standard_normal(N), 0) + 0.05 * rng.standard_normal(M)
    (tmp_path / "y.bin").write_bytes(y.tobytes())
    (tmp_path / "A_rows.bin").write_bytes(np.ascontiguousarray(A).tobytes())      # row-major
    (tmp_path / "A_cols.bin").write_bytes(np.asfortranarray(A).tobytes(order="F"))  # column-major
    kw = dict(maxiter=300, eps=1e-9, lambda_l1=0.05, lambda_tv=0.0, lambda_tsv=0.2, cinit=1.0)
    xl, rl, _, _ = sa.mfista_imaging_core_dft(y, A, NX, NY, nonneg_flag=1, **kw)
    for fname, extra in (("A_rows.bin", ["-t"]), ("A_cols.bin", [])):
        xo = tmp_path / f"x_{fname}"
        out = _run("mfista_imaging_dft", [M, N, tmp_path / "y.bin", tmp_path / fname, 0.05, 0, 0.2, 1.0, xo,
                                          "-rec", NX, "-nonneg", "-maxiter", 300, "-eps", 1e-9, *extra])
        x = df.read_vector(xo, N)
        e = np.linalg.norm(x - xl.ravel()) / np.linalg.norm(xl)
        print(f"  dft tool ({fname}) vs library relL2 {e:.2e}", flush=True)
        assert e <= 1e-12
        assert _report_value(out, "size of image:") == f"{NX} x {NY}"
        assert int(_report_value(out, "# of iterations:")) == rl["ITER"]
