"""Build libmfista_amd.so (HIP, gfx950) in-tree.

Every .hip file under csrc/ is compiled by hipcc in parallel, every .cpp
file (host-only code) by g++, and all are linked into
``sparseimaging_amd/lib/libmfista_amd.so``.  Packed-FP32 instructions are
disabled for all device code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
The command-line tools under cli/ are then linked against the library into
``sparseimaging_amd/bin/`` (run path ``$ORIGIN/../lib``).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libmfista_amd.so")
CLIDIR = os.path.join(HERE, "cli")
BINDIR = os.path.join(HERE, "bin")
TOOLS = ("mfista_imaging_nufft", "mfista_imaging_fft", "mfista_imaging_dft")
CXX = os.environ.get("CXX", "g++")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HOST_CFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFISTA_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    if src.endswith(".cpp"):
        cmd = [CXX, *HOST_CFLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,--no-undefined", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libmfista_amd.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_tools(verbose)
    return LIB


def build_tools(verbose: bool = True) -> None:
    """Links the command-line tools against the built library."""
    os.makedirs(BINDIR, exist_ok=True)
    for name in TOOLS:
        out = os.path.join(BINDIR, name)
        cmd = [CXX, *HOST_CFLAGS, os.path.join(CLIDIR, name + ".cpp"), "-o", out + ".tmp", f"-L{LIBDIR}",
               "-lmfista_amd", "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"build of {name} failed")
        os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {', '.join(TOOLS)} in {BINDIR}")


if __name__ == "__main__":
    build()
