"""File formats and command-line tools (CPU only: no solve runs here).

The reference's two example data files are pinned byte for byte: writing
their arrays back with the reference's 6-digit %e form must give files with
the recorded size and SHA-256 (tests/golden/ref_data_files.json), and reading
those files must give the arrays back exactly."""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import sparseimaging_amd as sa
from sparseimaging_amd import datafiles as df

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")
IO_HEADER = os.path.join(ROOT, "include", "mfista_io.h")


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def test_io_header_symbols_exported():
    import re

    src = re.sub(r"/\*.*?\*/", "", open(IO_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"^\s*(?:unsigned\s+)?(?:void|int|long)\s*\*?\s*([A-Za-z_]\w*)\s*\(", src,
                                  flags=re.M)))
    assert sorted(df.IO_EXPORTED) == names
    for n in names:
        assert hasattr(sa.lib(), n), n


def test_reference_nufft_file_round_trip(tmp_path):
    ref = json.load(open(os.path.join(G, "ref_data_files.json")))["nufft_data.txt"]
    d = np.load(os.path.join(G, "ref_nufft_data.npz"))
    p = tmp_path / "nufft_data.txt"
    df.write_nufft_data(p, d["u"], d["v"], d["vis_r"], d["vis_i"], d["std"], int(d["NX"]), int(d["NY"]), digits=6)
    assert os.path.getsize(p) == ref["bytes"]
    assert _sha(p) == ref["sha256"]
    r = df.read_nufft_data(p)
    assert (r["M"], r["NX"], r["NY"]) == (d["u"].size, int(d["NX"]), int(d["NY"]))
    for k in ("u", "v", "vis_r", "vis_i", "std"):
        assert np.array_equal(r[k], d[k]), k


def test_reference_fft_file_round_trip(tmp_path):
    ref = json.load(open(os.path.join(G, "ref_data_files.json")))["fft_data.txt"]
    d = np.load(os.path.join(G, "ref_fft_data.npz"))
    p = tmp_path / "fft_data.txt"
    df.write_fft_data(p, d["u"], d["v"], d["y_r"], d["y_i"], d["noise"], int(d["NX"]), int(d["NY"]), digits=6)
    assert os.path.getsize(p) == ref["bytes"]
    assert _sha(p) == ref["sha256"]
    r = df.read_fft_data(p)
    assert np.array_equal(r["u_idx"], d["u"]) and np.array_equal(r["v_idx"], d["v"])
    for k, kd in (("y_r", "y_r"), ("y_i", "y_i"), ("noise", "noise")):
        assert np.array_equal(r[k], d[kd]), k


def test_full_precision_round_trip_and_tolerant_header(tmp_path):
    rng = np.random.default_rng(3)
    a = [rng.standard_normal(37) * 10.0 ** rng.integers(-300, 300, 37) for _ in range(5)]
    p = tmp_path / "x.txt"
    df.write_nufft_data(p, *a, 8, 6)
    r = df.read_nufft_data(p)
    for k, ak in zip(("u", "v", "vis_r", "vis_i", "std"), a):
        assert np.array_equal(r[k], ak)
    # header spacing as fscanf's whitespace directives accept it, CRLF rows, no title line
    q = tmp_path / "loose.txt"
    q.write_bytes(b"M=2\r\nNX =4\r\nNY= 6\r\n\r\n1.5,2.5 , 3, 4,5\r\n-1e-3, 2E+2,0,0,1\r\n")
    r = df.read_nufft_data(q)
    assert (r["M"], r["NX"], r["NY"]) == (2, 4, 6)
    assert np.array_equal(r["v"], [2.5, 200.0]) and np.array_equal(r["std"], [5.0, 1.0])


def test_data_file_errors(tmp_path):
    with pytest.raises(df.MfistaIOError) as e:
        df.read_nufft_data(tmp_path / "missing.txt")
    assert e.value.status == df.IO_E_OPEN
    bad = tmp_path / "bad.txt"
    bad.write_text("M  = 3\nNX = 4\nNY = 4\n\nu_rad, v_rad, vis_r, vis_i, noise_std_dev\n\n1,2,3,4,5\n1,2,x,4,5\n")
    with pytest.raises(df.MfistaIOError) as e:
        df.read_nufft_data(bad)
    assert e.value.status == df.IO_E_ROW
    short = tmp_path / "short.txt"
    short.write_text("M  = 3\nNX = 4\nNY = 4\n\nu, v, y_r, y_i, noise_std_dev\n\n1, 2, 3, 4, 5\n")
    with pytest.raises(df.MfistaIOError) as e:
        df.read_fft_data(short)
    assert e.value.status == df.IO_E_ROW
    nohdr = tmp_path / "nohdr.txt"
    nohdr.write_text("1, 2, 3, 4, 5\n")
    with pytest.raises(df.MfistaIOError) as e:
        df.read_fft_data(nohdr)
    assert e.value.status == df.IO_E_HEADER


def test_binary_vectors_and_transpose(tmp_path):
    x = np.arange(12, dtype=np.float64) * 0.5
    p = tmp_path / "x.bin"
    df.write_vector(p, x)
    assert p.read_bytes() == x.tobytes()
    assert np.array_equal(df.read_vector(p, 12), x)
    assert np.array_equal(df.read_vector(p, 20), x)  # a short file returns what it holds
    f = tmp_path / "box.bin"
    f.write_bytes(np.array([1, 0, 1], np.float32).tobytes())
    assert np.array_equal(df.read_vector(f, 3, np.float32), [1, 0, 1])
    a = np.arange(6.0).reshape(2, 3)
    assert np.array_equal(df.transpose_matrix(a), a.T)
    with pytest.raises(df.MfistaIOError):
        df.read_vector(tmp_path / "none.bin", 3)


def test_report_format(tmp_path):
    """show_io_fnames / show_result write the report blocks of c/mfista_io.c:141-215."""
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    r = sa.RESULT()
    r.M, r.N, r.NX, r.NY, r.N_active, r.maxiter, r.ITER, r.nonneg = 100, 64, 8, 8, 12, 500, 321, 1
    r.lambda_l1, r.lambda_tsv, r.finalcost, r.comp_time, r.Lip_const = 1.0, 10.0, 12.5, 0.25, 3e4
    r.sq_error, r.mean_sq_error, r.l1cost, r.tsvcost = 20.0, 0.2, 3.0, 0.05
    path = tmp_path / "report.txt"
    fp = libc.fopen(os.fsencode(path), b"w")

    class IO(ctypes.Structure):
        _fields_ = [("fft", ctypes.c_uint)] + [(n, ctypes.c_char_p) for n in
                                                ("fft_fname", "v_fname", "A_fname", "in_fname", "out_fname")]

    io = IO(1, b"data.txt", None, None, None, b"x.out")
    L = df._L()
    L.show_io_fnames(fp, b"prog", ctypes.addressof(io))
    L.show_result(fp, b"prog", ctypes.addressof(r))
    libc.fclose(fp)
    text = path.read_text()
    expect = [
        "IO files of prog.", " FFTW file:              data.txt", " x is saved to:          x.out",
        "Output of prog.", " size of input vector:   100", " size of output vector:  64",
        " size of image:          8 x 8", " x is a nonnegative vector.", " Lambda_1:               1.000000e+00",
        " Lambda_TSV:             1.000000e+01", " MAXITER:                500", " # of iterations:        321",
        " cost:                   1.250000e+01", " computaion time[sec]:   2.500000e-01",
        " Est. Lipschitzs const:  3.000000e+04", " # of nonzero pixels:    12",
        " Squared Error (SE):     2.000000e+01", " Mean SE:                2.000000e-01",
        " L1 cost:                3.000000e+00", " TSV cost:               5.000000e-02",
    ]
    lines = text.splitlines()
    for e in expect:
        assert e in lines, e
    assert "Lambda_TV" not in text and "TV cost" not in text and "x was initialized" not in text


@pytest.mark.parametrize("tool", df.TOOLS)
def test_tools_usage_and_missing_file(tool, tmp_path):
    exe = df.tool_path(tool)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "<X outfile>" in r.stdout
    if tool == "mfista_imaging_dft":
        args = ["4", "4", str(tmp_path / "y.bin"), str(tmp_path / "A.bin"), "1", "0", "0", "1e4", str(tmp_path / "x")]
    else:
        args = [str(tmp_path / "none.txt"), "1", "0", "0", "1e4", str(tmp_path / "x")]
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    assert r.returncode == 1
    assert "Can't fopen(r)" in r.stderr


def test_tool_malformed_file_exits_zero(tmp_path):
    """c/mfista_imaging_nufft.c:159-164: a row that does not parse prints
    "cannot read data." and exits with status 0, before any GPU work."""
    bad = tmp_path / "bad.txt"
    bad.write_text("M  = 2\nNX = 4\nNY = 4\n\nu_rad, v_rad, vis_r, vis_i, noise_std_dev\n\n1,2,3,4,5\n")
    r = subprocess.run([df.tool_path("mfista_imaging_nufft"), str(bad), "1", "0", "0", "1e4", str(tmp_path / "x")],
                       capture_output=True, text=True)
    assert r.returncode == 0 and "cannot read data." in r.stdout
    assert not (tmp_path / "x").exists()
