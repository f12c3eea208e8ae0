"""FP64 CPU oracle for the MFISTA imaging path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``liboracle_mfista.so`` (built from ``mfista_oracle.c`` by
``oracle/Makefile``).  The oracle restates the reference C/C++ library
(ikeda46/sparseimaging ``c/`` and ``c++/``) in plain FP64 C; every C function
cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``sparseimaging_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_mfista.so")
# the same source built with OpenMP: the all-cores CPU baseline (timing only;
# the checker is the serial build)
_LIB_OMP_PATH = os.path.join(_HERE, "liboracle_mfista_omp.so")

PRESET_C24 = 0
PRESET_CXX12 = 1
MAXITER, MINITER, FGPITER, TD, ETA = 50000, 100, 100, 50, 1.1


class Result(C.Structure):
    """Same field order as struct RESULT (c/mfista.h:45-66)."""

    _fields_ = [(n, C.c_int) for n in ("M", "N", "NX", "NY", "N_active", "maxiter", "ITER", "nonneg")] + [
        (n, C.c_double)
        for n in (
            "lambda_l1", "lambda_tv", "lambda_tsv", "sq_error", "mean_sq_error", "l1cost",
            "tvcost", "tsvcost", "finalcost", "comp_time",
        )
    ] + [("residual", C.POINTER(C.c_double)), ("Lip_const", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "residual"}


class Opts(C.Structure):
    _fields_ = [
        ("preset", C.c_int),
        ("fgp_iter", C.c_int),
        ("rel_tol", C.c_double),
        ("verbose", C.c_int),
        ("cost_trace", C.POINTER(C.c_double)),
        ("max_trials_out", C.c_int),
    ]


_lib = None
_lib_omp = None


def build(force: bool = False) -> str:
    if force or not (os.path.exists(_LIB_PATH) and os.path.exists(_LIB_OMP_PATH)):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)
    return _LIB_PATH


def lib(omp: bool = False):
    """The serial oracle (the checker), or with omp=True its OpenMP build."""
    global _lib, _lib_omp
    if omp:
        if _lib_omp is None:
            if not os.path.exists(_LIB_OMP_PATH):
                build()
            _lib_omp = C.CDLL(_LIB_OMP_PATH)
            _declare(_lib_omp)
        return _lib_omp
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


def _declare(L):
    L.or_fft_c2c.argtypes = [C.c_int, _dp, C.c_int]
    L.or_fft2_r2c.argtypes = [C.c_int, C.c_int, _dp, _dp]
    L.or_fft2_c2r.argtypes = [C.c_int, C.c_int, _dp, _dp]
    L.or_TSV.restype = C.c_double
    L.or_TSV.argtypes = [C.c_int, C.c_int, _dp]
    L.or_d_TSV.argtypes = [C.c_int, C.c_int, _dp, _dp]
    L.or_TV.restype = C.c_double
    L.or_TV.argtypes = [C.c_int, C.c_int, _dp]
    L.or_Lx.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_Lpq.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_P_pqiso.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    L.or_calc_Q_part.restype = C.c_double
    L.or_calc_Q_part.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
    L.or_soft_threshold_box.argtypes = [_dp, C.c_int, C.c_double, _dp, C.c_int, _fp]
    L.or_soft_threshold_nonneg_box.argtypes = [_dp, C.c_int, C.c_double, _dp, C.c_int, _fp]
    L.or_FGP_L1_box.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_int, _dp, C.c_int, _fp]
    L.or_FGP_nonneg_box.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_int, _dp, C.c_int, _fp]
    L.or_nufft_forward.restype = C.c_int
    L.or_nufft_forward.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_nufft_adjoint.restype = C.c_int
    L.or_nufft_adjoint.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_fft_F_grad.restype = C.c_double
    L.or_fft_F_grad.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.or_imaging_core_nufft.restype = C.c_int
    L.or_imaging_core_nufft.argtypes = [
        _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
        C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int, _fp, C.POINTER(Result),
        C.POINTER(Opts),
    ]
    L.or_imaging_core_fft.restype = C.c_int
    L.or_imaging_core_fft.argtypes = [
        _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
        C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int, _fp, C.POINTER(Result),
        C.POINTER(Opts),
    ]
    L.or_imaging_core_dft.restype = C.c_int
    L.or_imaging_core_dft.argtypes = [
        _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
        C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int, _fp, C.POINTER(Result), C.POINTER(Opts),
    ]
    L.or_result_nufft.restype = C.c_int
    L.or_result_nufft.argtypes = [
        C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
        C.c_double, _dp, C.POINTER(Result),
    ]
    L.or_free.argtypes = [C.c_void_p]


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _f(a, n):
    if a is None:
        a = np.ones(n, dtype=np.float32)
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


# ---------------------------------------------------------------- operators
def fft2_r2c(x):
    x, px = _d(x)
    n0, n1 = x.shape
    out = np.zeros((n0, n1 // 2 + 1), dtype=np.complex128)
    lib().or_fft2_r2c(n0, n1, px, out.ctypes.data_as(_dp))
    return out


def fft2_c2r(h, n1):
    h = np.ascontiguousarray(h, dtype=np.complex128)
    n0 = h.shape[0]
    out = np.zeros((n0, n1), dtype=np.float64)
    lib().or_fft2_c2r(n0, n1, h.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def fft_c2c(z, sign):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().or_fft_c2c(z.size, z.ctypes.data_as(_dp), sign)
    return z


def tsv(x):
    x, px = _d(x)
    NX, NY = x.shape[1], x.shape[0]  # flat index NX*j+i, i.e. the array viewed as NY rows of NX
    return lib().or_TSV(NX, NY, px)


def tsv_flat(x, NX, NY):
    x, px = _d(x)
    return lib().or_TSV(NX, NY, px)


def d_tsv_flat(x, NX, NY):
    x, px = _d(x)
    d = np.zeros(NX * NY)
    lib().or_d_TSV(NX, NY, px, d.ctypes.data_as(_dp))
    return d


def tv_flat(x, NX, NY):
    x, px = _d(x)
    return lib().or_TV(NX, NY, px)


def Lx_flat(x, NX, NY):
    x, px = _d(x)
    p = np.zeros((NX - 1) * NY)
    q = np.zeros(NX * (NY - 1))
    lib().or_Lx(NX, NY, px, p.ctypes.data_as(_dp), q.ctypes.data_as(_dp))
    return p, q


def Lpq_flat(p, q, NX, NY):
    p, pp = _d(p)
    q, pq = _d(q)
    x = np.zeros(NX * NY)
    lib().or_Lpq(NX, NY, pp, pq, x.ctypes.data_as(_dp))
    return x


def P_pqiso_flat(p, q, NX, NY):
    p, pp = _d(p)
    q, pq = _d(q)
    r = np.zeros_like(p)
    s = np.zeros_like(q)
    lib().or_P_pqiso(NX, NY, pp, pq, r.ctypes.data_as(_dp), s.ctypes.data_as(_dp))
    return r, s


def calc_Q_part(x1, x2, c, g):
    x1, p1 = _d(x1)
    x2, p2 = _d(x2)
    g, pg = _d(g)
    buf = np.zeros_like(x1)
    return lib().or_calc_Q_part(x1.size, p1, p2, c, pg, buf.ctypes.data_as(_dp))


def soft_threshold(v, eta, nonneg=False, box=None):
    v, pv = _d(v)
    out = np.zeros_like(v)
    bflag = 0 if box is None else 1
    b, pb = _f(box, v.size)
    f = lib().or_soft_threshold_nonneg_box if nonneg else lib().or_soft_threshold_box
    f(pv, v.size, eta, out.ctypes.data_as(_dp), bflag, pb)
    return out


def fgp(b, NX, NY, lambda_l1, lambda_tv, iters, nonneg=False, box=None):
    b, pb = _d(b)
    x = np.zeros(NX * NY)
    bflag = 0 if box is None else 1
    bx, pbx = _f(box, NX * NY)
    if nonneg:
        lib().or_FGP_nonneg_box(NX, NY, pb, lambda_tv, iters, x.ctypes.data_as(_dp), bflag, pbx)
    else:
        lib().or_FGP_L1_box(NX, NY, pb, lambda_l1, lambda_tv, iters, x.ctypes.data_as(_dp), bflag, pbx)
    return x


def nufft_forward(u, v, x, preset=PRESET_C24):
    """A.x for an NX x NY image x (row-major x[i*NY+j]); returns complex model visibilities."""
    u, pu = _d(u)
    v, pv = _d(v)
    x, px = _d(x)
    NX, NY = x.shape
    yr = np.zeros(u.size)
    yi = np.zeros(u.size)
    rc = lib().or_nufft_forward(preset, pu, pv, u.size, NX, NY, px, yr.ctypes.data_as(_dp), yi.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("FOV violation")
    return yr + 1j * yi


def nufft_adjoint(u, v, y, NX, NY, preset=PRESET_C24):
    """Re(A^H y) as NUFFT2d1 computes it; returns an NX x NY image."""
    u, pu = _d(u)
    v, pv = _d(v)
    yr, pyr = _d(np.real(y))
    yi, pyi = _d(np.imag(y))
    x = np.zeros((NX, NY))
    rc = lib().or_nufft_adjoint(preset, pu, pv, u.size, NX, NY, pyr, pyi, x.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("FOV violation")
    return x


def fft_F_grad(u_idx, v_idx, y_r, y_i, noise, x):
    u_idx, pu = _i(u_idx)
    v_idx, pv = _i(v_idx)
    y_r, pyr = _d(y_r)
    y_i, pyi = _d(y_i)
    noise, pn = _d(noise)
    x, px = _d(x)
    NX, NY = x.shape
    g = np.zeros((NX, NY))
    F = lib().or_fft_F_grad(u_idx.size, NX, NY, pu, pv, pyr, pyi, pn, px, g.ctypes.data_as(_dp))
    return F, g


# ---------------------------------------------------------------- solvers
def _opts(preset, fgp_iter, rel_tol, maxiter, trace):
    o = Opts()
    o.preset = preset
    o.fgp_iter = fgp_iter
    o.rel_tol = rel_tol
    o.verbose = 0
    buf = None
    if trace:
        buf = np.zeros(max(maxiter, 1))
        o.cost_trace = buf.ctypes.data_as(_dp)
    return o, buf


def _finish(rc, res, xout, opts, buf):
    out = {"rc": rc, "x": xout, "result": res.as_dict(), "max_trials": opts.max_trials_out}
    if buf is not None and rc == 0:
        out["cost"] = buf[: max(res.ITER, 0)].copy()
    return out


def imaging_core_nufft(u, v, vis_r, vis_i, vis_std, NX, NY, maxiter, eps, lambda_l1, lambda_tv,
                       lambda_tsv, cinit, xinit=None, nonneg=1, cl_box=None, preset=PRESET_C24,
                       fgp_iter=FGPITER, rel_tol=0.0, trace=False, omp=False):
    u, pu = _d(u)
    v, pv = _d(v)
    vr, pvr = _d(vis_r)
    vi, pvi = _d(vis_i)
    vs, pvs = _d(vis_std)
    x0 = np.zeros(NX * NY) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(NX * NY)
    bflag = 0 if cl_box is None else 1
    bx, pbx = _f(cl_box, NX * NY)
    res = Result()
    o, buf = _opts(preset, fgp_iter, rel_tol, maxiter, trace)
    rc = lib(omp).or_imaging_core_nufft(pu, pv, pvr, pvi, pvs, u.size, NX, NY, maxiter, eps, lambda_l1,
                                     lambda_tv, lambda_tsv, cinit, px0, xout.ctypes.data_as(_dp),
                                     nonneg, bflag, pbx, C.byref(res), C.byref(o))
    return _finish(rc, res, xout.reshape(NX, NY), o, buf)


def result_nufft(u, v, vis_r, vis_i, vis_std, x, lambda_l1, lambda_tv, lambda_tsv, preset=PRESET_C24):
    u, pu = _d(u)
    v, pv = _d(v)
    vr, pvr = _d(vis_r)
    vi, pvi = _d(vis_i)
    vs, pvs = _d(vis_std)
    x, px = _d(x)
    NX, NY = x.shape
    res = Result()
    rc = lib().or_result_nufft(preset, pu, pv, pvr, pvi, pvs, u.size, NX, NY, lambda_l1, lambda_tv,
                               lambda_tsv, px, C.byref(res))
    if rc != 0:
        raise ValueError("FOV violation")
    return res.as_dict()


def imaging_core_fft(u_idx, v_idx, y_r, y_i, noise, NX, NY, maxiter, eps, lambda_l1, lambda_tv,
                     lambda_tsv, cinit, xinit=None, nonneg=1, cl_box=None, fgp_iter=FGPITER,
                     rel_tol=0.0, trace=False):
    ui, pui = _i(u_idx)
    vi_, pvi_ = _i(v_idx)
    yr, pyr = _d(y_r)
    yi, pyi = _d(y_i)
    ns, pns = _d(noise)
    x0 = np.zeros(NX * NY) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(NX * NY)
    bflag = 0 if cl_box is None else 1
    bx, pbx = _f(cl_box, NX * NY)
    res = Result()
    o, buf = _opts(PRESET_C24, fgp_iter, rel_tol, maxiter, trace)
    rc = lib().or_imaging_core_fft(pui, pvi_, pyr, pyi, pns, ui.size, NX, NY, maxiter, eps, lambda_l1,
                                   lambda_tv, lambda_tsv, cinit, px0, xout.ctypes.data_as(_dp), nonneg,
                                   bflag, pbx, C.byref(res), C.byref(o))
    return _finish(rc, res, xout.reshape(NX, NY), o, buf)


def imaging_core_dft(y, A, NX, NY, maxiter, eps, lambda_l1, lambda_tv, lambda_tsv, cinit, xinit=None,
                     nonneg=1, cl_box=None, fgp_iter=FGPITER, rel_tol=0.0, trace=False):
    """A is M x N (numpy, any order); passed column-major as the reference expects."""
    A = np.asarray(A, dtype=np.float64)
    M, N = A.shape
    Acm = np.asfortranarray(A)
    pA = Acm.ctypes.data_as(_dp)
    y, py = _d(y)
    x0 = np.zeros(N) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(N)
    bflag = 0 if cl_box is None else 1
    bx, pbx = _f(cl_box, N)
    res = Result()
    o, buf = _opts(PRESET_C24, fgp_iter, rel_tol, maxiter, trace)
    rc = lib().or_imaging_core_dft(py, pA, M, N, NX, NY, maxiter, eps, lambda_l1, lambda_tv, lambda_tsv,
                                   cinit, px0, xout.ctypes.data_as(_dp), nonneg, bflag, pbx,
                                   C.byref(res), C.byref(o))
    out = _finish(rc, res, xout.reshape(NX, NY), o, buf)
    if rc == 0 and res.residual:
        out["residual"] = np.ctypeslib.as_array(res.residual, shape=(M,)).copy()
        lib().or_free(C.cast(res.residual, C.c_void_p))
    return out
