"""Data files of the imaging tools, through the library's own readers and
writers (include/mfista_io.h; host code, no GPU needed).

- NUFFT text file (c/mfista_imaging_nufft.c:140-170): header ``M  = ..``,
  ``NX = ..``, ``NY = ..``, a title line, then ``u_rad, v_rad, vis_r, vis_i,
  noise_std_dev`` rows.
- FFT text file (c/mfista_imaging_fft.c:153-183): the same header, then
  ``u, v, y_r, y_i, noise_std_dev`` rows with integer cells.
- Binary vectors (c/mfista_io.c:27-114): raw native-endian float64 images and
  DFT data, float32 CLEAN boxes, int32 index vectors.

Command-line tools of the same names as the reference's sit in
``sparseimaging_amd/bin/`` (``tool_path(name)``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import lib as _lib

_HERE = os.path.dirname(os.path.abspath(__file__))
BIN_DIR = os.path.join(_HERE, "bin")
TOOLS = ("mfista_imaging_nufft", "mfista_imaging_fft", "mfista_imaging_dft")

IO_OK, IO_E_OPEN, IO_E_HEADER, IO_E_ROW, IO_E_ARG, IO_E_NOMEM = 0, -1, -2, -3, -4, -5
IO_EXPORTED = [
    "read_int_vector", "read_f_vector", "read_V_vector", "read_A_matrix", "write_X_vector",
    "transpose_matrix", "show_io_fnames", "show_result", "mfista_read_nufft_data", "mfista_read_fft_data",
    "mfista_write_nufft_data", "mfista_write_fft_data", "mfista_io_free",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_declared = False


class MfistaIOError(RuntimeError):
    def __init__(self, status, path):
        msg = {IO_E_OPEN: "cannot open", IO_E_HEADER: "header does not parse", IO_E_ROW: "cannot read data",
               IO_E_ARG: "invalid argument", IO_E_NOMEM: "out of memory"}.get(status, "error")
        super().__init__(f"{path}: {msg} (status {status})")
        self.status = status


def _L():
    global _declared
    L = _lib()
    if not _declared:
        i, s = C.c_int, C.c_char_p
        L.mfista_read_nufft_data.argtypes = [s, _ip, _ip, _ip] + [C.POINTER(_dp)] * 5
        L.mfista_read_nufft_data.restype = i
        L.mfista_read_fft_data.argtypes = [s, _ip, _ip, _ip, C.POINTER(_ip), C.POINTER(_ip)] + [C.POINTER(_dp)] * 3
        L.mfista_read_fft_data.restype = i
        L.mfista_write_nufft_data.argtypes = [s, i, i, i] + [_dp] * 5 + [i]
        L.mfista_write_nufft_data.restype = i
        L.mfista_write_fft_data.argtypes = [s, i, i, i, _ip, _ip, _dp, _dp, _dp, i]
        L.mfista_write_fft_data.restype = i
        L.mfista_io_free.argtypes = [C.c_void_p]
        L.mfista_io_free.restype = None
        L.read_V_vector.argtypes = [s, i, _dp]
        L.read_V_vector.restype = i
        L.read_f_vector.argtypes = [s, i, C.POINTER(C.c_float)]
        L.read_f_vector.restype = i
        L.read_int_vector.argtypes = [s, i, _ip]
        L.read_int_vector.restype = i
        L.read_A_matrix.argtypes = [s, i, i, _dp]
        L.read_A_matrix.restype = C.c_ulong
        L.write_X_vector.argtypes = [s, i, _dp]
        L.write_X_vector.restype = i
        L.transpose_matrix.argtypes = [_dp, i, i]
        L.transpose_matrix.restype = None
        L.show_io_fnames.argtypes = [C.c_void_p, s, C.c_void_p]
        L.show_io_fnames.restype = None
        L.show_result.argtypes = [C.c_void_p, s, C.c_void_p]
        L.show_result.restype = None
        _declared = True
    return L


def _take(p, n, dtype):
    out = np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True) if n > 0 else np.empty(0, dtype)
    _L().mfista_io_free(C.cast(p, C.c_void_p))
    return out


def read_nufft_data(path):
    """dict(u, v, vis_r, vis_i, std, M, NX, NY) of a NUFFT data file."""
    M, NX, NY = C.c_int(), C.c_int(), C.c_int()
    ptrs = [_dp() for _ in range(5)]
    st = _L().mfista_read_nufft_data(os.fsencode(path), C.byref(M), C.byref(NX), C.byref(NY),
                                     *[C.byref(p) for p in ptrs])
    if st != IO_OK:
        raise MfistaIOError(st, path)
    names = ("u", "v", "vis_r", "vis_i", "std")
    out = {n: _take(p, M.value, np.float64) for n, p in zip(names, ptrs)}
    out.update(M=M.value, NX=NX.value, NY=NY.value)
    return out


def read_fft_data(path):
    """dict(u_idx, v_idx, y_r, y_i, noise, M, NX, NY) of an FFT data file."""
    M, NX, NY = C.c_int(), C.c_int(), C.c_int()
    iu, iv = _ip(), _ip()
    ptrs = [_dp() for _ in range(3)]
    st = _L().mfista_read_fft_data(os.fsencode(path), C.byref(M), C.byref(NX), C.byref(NY), C.byref(iu),
                                   C.byref(iv), *[C.byref(p) for p in ptrs])
    if st != IO_OK:
        raise MfistaIOError(st, path)
    out = {"u_idx": _take(iu, M.value, np.int32), "v_idx": _take(iv, M.value, np.int32)}
    out.update({n: _take(p, M.value, np.float64) for n, p in zip(("y_r", "y_i", "noise"), ptrs)})
    out.update(M=M.value, NX=NX.value, NY=NY.value)
    return out


def _dv(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def write_nufft_data(path, u, v, vis_r, vis_i, std, NX, NY, digits=16):
    """Writes a NUFFT data file; digits=6 gives the reference files' %e form,
    16 (default) round-trips every double."""
    arrs = [_dv(a) for a in (u, v, vis_r, vis_i, std)]
    M = arrs[0][0].size
    if any(a.size != M for a, _ in arrs):
        raise ValueError("u, v, vis_r, vis_i and std must have the same length")
    st = _L().mfista_write_nufft_data(os.fsencode(path), M, int(NX), int(NY), *[p for _, p in arrs], int(digits))
    if st != IO_OK:
        raise MfistaIOError(st, path)


def write_fft_data(path, u_idx, v_idx, y_r, y_i, noise, NX, NY, digits=16):
    iu = np.ascontiguousarray(u_idx, dtype=np.int32)
    iv = np.ascontiguousarray(v_idx, dtype=np.int32)
    arrs = [_dv(a) for a in (y_r, y_i, noise)]
    M = iu.size
    if iv.size != M or any(a.size != M for a, _ in arrs):
        raise ValueError("all columns must have the same length")
    st = _L().mfista_write_fft_data(os.fsencode(path), M, int(NX), int(NY), iu.ctypes.data_as(_ip),
                                    iv.ctypes.data_as(_ip), *[p for _, p in arrs], int(digits))
    if st != IO_OK:
        raise MfistaIOError(st, path)


def read_vector(path, n, dtype=np.float64):
    """read_V_vector / read_f_vector / read_int_vector: up to n values; returns
    the values read (fewer if the file is short)."""
    dtype = np.dtype(dtype)
    buf = np.zeros(int(n), dtype=dtype)
    L = _L()
    fn = {np.dtype(np.float64): (L.read_V_vector, _dp), np.dtype(np.float32): (L.read_f_vector, C.POINTER(C.c_float)),
          np.dtype(np.int32): (L.read_int_vector, _ip)}[dtype]
    got = fn[0](os.fsencode(path), int(n), buf.ctypes.data_as(fn[1]))
    if got < 0:
        raise MfistaIOError(IO_E_OPEN, path)
    return buf[:got]


def write_vector(path, x):
    """write_X_vector: raw float64."""
    a, p = _dv(np.ravel(x))
    n = _L().write_X_vector(os.fsencode(path), a.size, p)
    if n != a.size:
        raise MfistaIOError(IO_E_OPEN, path)


def transpose_matrix(a):
    """In-place row-major (h, w) -> row-major (w, h) on a copy; returns it."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    h, w = a.shape
    _L().transpose_matrix(a.ctypes.data_as(_dp), h, w)
    return a.reshape(w, h)


def tool_path(name):
    """Path of a built command-line tool (mfista_imaging_nufft / _fft / _dft)."""
    if name not in TOOLS:
        raise ValueError(name)
    p = os.path.join(BIN_DIR, name)
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: run __graft_entry__.build() first")
    return p
