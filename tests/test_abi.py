"""The C ABI (CPU only, no compute): the library loads, exports every function
include/mfista.h declares, and struct RESULT has the reference layout
(c/mfista.h:45-66: 8 ints, 10 doubles, a pointer and a double; 128 bytes)."""
import ctypes
import os
import re
import subprocess

import pytest

import sparseimaging_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfista.h")
LEARN_HEADER = os.path.join(ROOT, "include", "mfista_learn.h")


def _declared_functions(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|char|long|double)\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = sa.lib()
    names = _declared_functions()
    assert "mfista_imaging_core_nufft" in names and "y2x_nufft" in names
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(sa.EXPORTED) == names


def test_learn_header_symbols_exported():
    lib = sa.lib()
    names = _declared_functions(LEARN_HEADER)
    assert "unet_conv3x3_fwd" in names and "unet_adam" in names
    for n in names:
        assert hasattr(lib, n), n


def test_reference_symbols_present():
    for n in ("mfista_imaging_core_nufft", "mfista_imaging_core_fft", "mfista_imaging_core_dft",
              "x2y_nufft", "y2x_nufft"):
        assert hasattr(sa.lib(), n)


def test_result_layout_matches_reference(tmp_path):
    prog = tmp_path / "layout.c"
    prog.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "mfista.h"\n'
        "int main(void){struct RESULT r; (void)r;\n"
        'printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(struct RESULT), offsetof(struct RESULT, nonneg),'
        " offsetof(struct RESULT, lambda_l1), offsetof(struct RESULT, comp_time),"
        " offsetof(struct RESULT, residual), offsetof(struct RESULT, Lip_const)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out] == [128, 28, 32, 104, 112, 120]
    assert ctypes.sizeof(sa.RESULT) == 128
    assert sa.RESULT.residual.offset == 112 and sa.RESULT.Lip_const.offset == 120


def test_batch_problem_and_opts_layout_match_ctypes(tmp_path):
    prog = tmp_path / "layout2.c"
    prog.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "mfista.h"\n'
        "int main(void){\n"
        'printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(mfista_nufft_problem), offsetof(mfista_nufft_problem, M),'
        " offsetof(mfista_nufft_problem, eps), offsetof(mfista_nufft_problem, xinit),"
        " offsetof(mfista_nufft_problem, cl_box), sizeof(mfista_opts)); return 0;}\n")
    exe = tmp_path / "layout2"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    P = sa.NufftProblem
    assert out == [ctypes.sizeof(P), P.M.offset, P.eps.offset, P.xinit.offset, P.cl_box.offset,
                   ctypes.sizeof(sa.Opts)]


def test_tune_struct_layout_matches_ctypes(tmp_path):
    prog = tmp_path / "layout3.c"
    prog.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "mfista.h"\n'
        "int main(void){\n"
        'printf("%zu %zu %zu %zu\\n", sizeof(mfista_tune), offsetof(mfista_tune, tile_x),'
        " offsetof(mfista_tune, paired), offsetof(mfista_tune, us_per_iter)); return 0;}\n")
    exe = tmp_path / "layout3"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    T = sa.Tune
    assert out == [ctypes.sizeof(T), T.tile_x.offset, T.paired.offset, T.us_per_iter.offset]


def test_tune_table_file_round_trip(tmp_path):
    """The autotuner's table file (no GPU work): a saved entry loads back, and
    lines with an unsupported tile or a bad field count are skipped."""
    f = tmp_path / "tune.txt"
    f.write_text("# header\n256 256 1 1 14 0 32 16 2 1 0 0 88.5\n256 256 1 1 14 3 24 16 2 1 0 0 1.0\n"
                 "256 256 1 1 14 4 16 16 2 1 7 0 1.0\n256 256 0 2 14 0 16 16 2 0 0 12 1.0\n1 2 3\n"
                 "128 128 1 1 13 0 16 16 1 1 0 0 2 70.25\n128 128 1 1 12 0 16 16 1 1 0 0 5 70.25\n")
    sa.tune_clear()
    assert sa.tune_load(str(f)) == 2  # a 13-field line (loop by rule) and a 14-field one
    t = sa.tune_table()
    t = sorted(t, key=lambda e: e["NX"], reverse=True)
    assert len(t) == 2 and t[0]["tile_x"] == 32 and t[0]["tile_y"] == 16 and t[0]["blocks_per_cu"] == 2
    assert t[0]["m_log2"] == 14 and t[0]["engine"] == sa.ENGINE_PERSISTENT and abs(t[0]["us_per_iter"] - 88.5) < 1e-9
    assert t[0]["loop"] == 0
    assert t[1]["NX"] == 128 and t[1]["loop"] == sa.LOOP_TOEPLITZ and abs(t[1]["us_per_iter"] - 70.25) < 1e-9
    g = tmp_path / "tune2.txt"
    sa.tune_save(str(g))
    sa.tune_clear()
    assert sa.tune_table() == []
    assert sa.tune_load(str(g)) == 2 and sorted(sa.tune_table(), key=lambda e: e["NX"], reverse=True) == t
    sa.tune_clear()


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "sparseimaging_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f


def test_python_constants_match_the_header():
    """The Python layer's engine, loop and preset values are the header's."""
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                             "mfista.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define MFISTA_(\w+) (-?\d+)\b", text, re.M)}
    for name in ("ENGINE_PERSISTENT", "ENGINE_STREAM", "LOOP_AUTO", "LOOP_GRIDDING", "LOOP_TOEPLITZ",
                 "PRESET_C24", "PRESET_CXX12"):
        assert getattr(sa, name) == defs[name], name
