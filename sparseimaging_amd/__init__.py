"""sparseimaging_amd -- MI355X-native MFISTA imaging solver (ctypes mirror of the C ABI).

The native library ``lib/libmfista_amd.so`` (HIP, gfx950) exports the
reference library's entry points (ikeda46/sparseimaging ``c/mfista.h:151-178``,
``c++/mfista.hpp:117-140``).  This module binds them with ctypes and adds
numpy-friendly wrappers whose argument names and meaning follow the reference
(``mfista_imaging_core_nufft(u_dx, v_dy, vis_r, vis_i, vis_std, M, NX, NY,
maxiter, eps, lambda_l1, lambda_tv, lambda_tsv, cinit, xinit, xout,
nonneg_flag, box_flag, cl_box, mfista_result)``).

There is no CPU fallback: importing works without a GPU (for ABI checks), but
every solve needs the HIP library and a visible gfx950 device and raises
otherwise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmfista_amd.so")

PRESET_C24 = 0
PRESET_CXX12 = 1
ENGINE_PERSISTENT = 0  # one cooperative launch per NUFFT solve
ENGINE_STREAM = 1      # one launch per kernel, device-side gating
LOOP_AUTO = 0          # the Toeplitz loop where it applies (or the tuned loop), else gridding
LOOP_GRIDDING = 1      # grid the residual every iteration, degrid every trial (the reference's structure)
LOOP_TOEPLITZ = 2      # the normal operator as an FFT convolution, no gridding in the loop (wherever it applies)
MAXITER, MINITER, FGPITER, TD, ETA, EPS = 50000, 100, 100, 50, 1.1, 1.0e-5

STATUS = {
    0: "ok", -1: "lambda_tv and lambda_tsv both positive", -2: "TV needs the C-24 preset",
    -3: "field of view too small for the (u,v) coverage", -4: "unsupported image size",
    -5: "invalid argument", -6: "nonneg_flag must be 0 or 1", -100: "HIP runtime error",
    -101: "cooperative launch refused (another process holds the device's cooperative-launch resources)",
}

EXPORTED = [
    "mfista_imaging_core_nufft", "mfista_imaging_core_fft", "mfista_imaging_core_dft",
    "x2y_nufft", "y2x_nufft", "mfista_default_opts", "mfista_imaging_core_nufft_ex",
    "mfista_nufft_solve_device", "mfista_nufft_forward", "mfista_nufft_adjoint",
    "mfista_amd_version", "mfista_amd_finalize", "mfista_profile_stages", "mfista_last_steps",
    "mfista_imaging_core_fft_ex", "mfista_imaging_core_dft_ex", "mfista_imaging_core_nufft_batch",
    "mfista_movie_nufft", "mfista_looe_dft", "compute_LOOE_L1", "compute_LOOE_L1_TSV",
    "mfista_nufft_solve_device_noalloc", "mfista_device_allocations",
    "mfista_dft_products", "mfista_imaging_core_dft_batch",
    "mfista_autotune_nufft", "mfista_autotune_nufft_batch", "mfista_tune_table", "mfista_tune_save",
    "mfista_tune_load", "mfista_tune_clear", "mfista_tune_runs",
]
STAGES = ["rowfft_fwd", "colfft_fwd", "degrid", "grid", "colfft_inv", "rowfft_c2r", "tsv_grad", "prox"]


class NufftProblem(C.Structure):
    """mfista_nufft_problem (include/mfista.h): one problem of a batch, host pointers."""

    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("u_dx", "v_dy", "vis_r", "vis_i", "vis_std")] + [
        (n, C.c_int) for n in ("M", "NX", "NY", "maxiter")] + [
        (n, C.c_double) for n in ("eps", "lambda_l1", "lambda_tv", "lambda_tsv", "cinit")] + [
        ("xinit", C.POINTER(C.c_double)), ("xout", C.POINTER(C.c_double)), ("nonneg_flag", C.c_int),
        ("box_flag", C.c_int), ("cl_box", C.POINTER(C.c_float)), ("target_objective", C.c_double)]


class RESULT(C.Structure):
    """struct RESULT, field for field c/mfista.h:45-66."""

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
        ("preset", C.c_int), ("fgp_iter", C.c_int), ("rel_tol", C.c_double), ("verbose", C.c_int),
        ("max_trials", C.c_int), ("solve_ms", C.c_double), ("total_trials", C.c_int),
        ("engine", C.c_int), ("kernel_ms", C.c_double), ("trial_rounds", C.c_int),
        ("single_trials", C.c_int), ("engine_used", C.c_int), ("deterministic", C.c_int),
        ("autotune", C.c_int), ("tune_iters", C.c_int), ("loop", C.c_int), ("loop_used", C.c_int),
    ]


class Tune(C.Structure):
    """mfista_tune: one entry of the autotuner's table (shape key and choice)."""
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int), ("preset", C.c_int), ("reg", C.c_int), ("m_log2", C.c_int),
        ("nprob_log2", C.c_int), ("tile_x", C.c_int), ("tile_y", C.c_int), ("blocks_per_cu", C.c_int),
        ("paired", C.c_int), ("engine", C.c_int), ("fgp_tw", C.c_int), ("us_per_iter", C.c_double),
        ("loop", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MfistaError(RuntimeError):
    def __init__(self, status):
        super().__init__(f"mfista_amd status {status}: {STATUS.get(status, 'error')}")
        self.status = status


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


def lib():
    """Load libmfista_amd.so; raises if it has not been built.

    torch (when installed) is imported first: its wheel ships its own HIP
    runtime, and a process must load only one.  Imported first, torch's runtime
    is the one the library binds to; loaded after the library has started the
    system runtime, torch cannot open the GPU ("No HIP GPUs are available")."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
        # free the library's streams and memory at interpreter exit, while the
        # HIP runtime is still up (Python's exit handlers run before the C
        # runtime's and the shared libraries' destructors)
        atexit.register(_lib.mfista_amd_finalize)
    return _lib


def _declare(L):
    i, d, u = C.c_int, C.c_double, C.c_uint
    L.mfista_imaging_core_nufft.restype = None
    L.mfista_imaging_core_nufft.argtypes = [_dp, _dp, _dp, _dp, _dp, i, i, i, i, d, d, d, d, d, _dp, _dp,
                                            i, i, _fp, C.POINTER(RESULT)]
    L.mfista_imaging_core_fft.restype = None
    L.mfista_imaging_core_fft.argtypes = [_ip, _ip, _dp, _dp, _dp, i, i, i, i, d, d, d, d, d, _dp, _dp,
                                          i, u, i, _fp, C.POINTER(RESULT)]
    L.mfista_imaging_core_dft.restype = None
    L.mfista_imaging_core_dft.argtypes = [_dp, _dp, _ip, _ip, i, i, i, d, d, d, d, d, _dp, _dp, i, i, i,
                                          _fp, C.POINTER(RESULT)]
    L.x2y_nufft.restype = None
    L.x2y_nufft.argtypes = [_dp, _dp, i, i, i, _dp, _dp, _dp]
    L.y2x_nufft.restype = None
    L.y2x_nufft.argtypes = [_dp, _dp, i, i, i, _dp, _dp, _dp]
    L.mfista_default_opts.restype = None
    L.mfista_default_opts.argtypes = [C.POINTER(Opts)]
    L.mfista_imaging_core_nufft_ex.restype = i
    L.mfista_imaging_core_nufft_ex.argtypes = [_dp, _dp, _dp, _dp, _dp, i, i, i, i, d, d, d, d, d, _dp,
                                               _dp, i, i, _fp, C.POINTER(RESULT), C.POINTER(Opts)]
    L.mfista_imaging_core_nufft_batch.restype = i
    L.mfista_imaging_core_nufft_batch.argtypes = [i, C.POINTER(NufftProblem), C.POINTER(RESULT), C.POINTER(Opts)]
    L.mfista_nufft_solve_device.restype = i
    L.mfista_nufft_solve_device.argtypes = [C.c_void_p] * 5 + [i, i, i, i, d, d, d, d, d] + [
        C.c_void_p, C.c_void_p, i, i, C.c_void_p, C.POINTER(RESULT), C.POINTER(Opts), C.c_void_p]
    L.mfista_nufft_solve_device_noalloc.restype = i
    L.mfista_nufft_solve_device_noalloc.argtypes = L.mfista_nufft_solve_device.argtypes
    L.mfista_device_allocations.restype = C.c_long
    L.mfista_device_allocations.argtypes = []
    L.mfista_nufft_forward.restype = i
    L.mfista_movie_nufft.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _dp, i, _fp, i, i, _dp, d, d, d, i, d, i, d,
                                     _dp, _dp, _dp, _dp, _dp, _ip, C.POINTER(Opts)]
    L.mfista_movie_nufft.restype = i
    L.mfista_nufft_forward.argtypes = [i, i, _dp, _dp, i, i, i, _dp, _dp, _dp]
    L.mfista_nufft_adjoint.restype = i
    L.mfista_nufft_adjoint.argtypes = [i, i, _dp, _dp, i, i, i, _dp, _dp, _dp]
    L.mfista_imaging_core_fft_ex.restype = i
    L.mfista_imaging_core_fft_ex.argtypes = [_ip, _ip, _dp, _dp, _dp, i, i, i, i, d, d, d, d, d, _dp, _dp,
                                             i, i, _fp, C.POINTER(RESULT), C.POINTER(Opts)]
    L.mfista_imaging_core_dft_ex.restype = i
    L.mfista_imaging_core_dft_ex.argtypes = [_dp, _dp, i, i, i, i, i, d, d, d, d, d, _dp, _dp, i, i, _fp,
                                             C.POINTER(RESULT), C.POINTER(Opts)]
    L.mfista_dft_products.restype = i
    L.mfista_dft_products.argtypes = [_dp, i, i, i, _dp, _dp, _dp, _dp]
    L.mfista_imaging_core_dft_batch.restype = i
    L.mfista_imaging_core_dft_batch.argtypes = [_dp, _dp, i, i, i, i, i, _dp, _dp, _dp, i, d, _dp, _dp, i, i,
                                                _fp, C.POINTER(RESULT), C.POINTER(Opts)]
    L.mfista_looe_dft.restype = i
    L.mfista_looe_dft.argtypes = [i, i, i, i, d, _dp, _dp, _dp, _dp, _dp, _ip, _ip]
    L.compute_LOOE_L1.restype = d
    L.compute_LOOE_L1.argtypes = [_ip, _ip, d, _dp, _dp, _dp, _dp, _dp, _dp]
    L.compute_LOOE_L1_TSV.restype = d
    L.compute_LOOE_L1_TSV.argtypes = [_ip, _ip, i, i, d, d, _dp, _dp, _dp, _dp, _dp, _dp]
    L.mfista_profile_stages.restype = i
    L.mfista_profile_stages.argtypes = [i, _dp, i]
    L.mfista_last_steps.restype = C.c_long
    L.mfista_last_steps.argtypes = []
    L.mfista_amd_version.restype = C.c_char_p
    L.mfista_amd_version.argtypes = []
    L.mfista_amd_finalize.restype = None
    L.mfista_amd_finalize.argtypes = []
    L.mfista_autotune_nufft.restype = i
    L.mfista_autotune_nufft.argtypes = [_dp, _dp, _dp, _dp, _dp, i, i, i, d, d, d, d, _dp, i, C.POINTER(Opts), i,
                                        C.POINTER(Tune), C.POINTER(Tune), i, _ip]
    L.mfista_autotune_nufft_batch.restype = i
    L.mfista_autotune_nufft_batch.argtypes = [i, C.POINTER(NufftProblem), C.POINTER(Opts), i, C.POINTER(Tune),
                                              C.POINTER(Tune), i, _ip]
    L.mfista_tune_table.restype = i
    L.mfista_tune_table.argtypes = [C.POINTER(Tune), i]
    L.mfista_tune_save.restype = i
    L.mfista_tune_save.argtypes = [C.c_char_p]
    L.mfista_tune_load.restype = i
    L.mfista_tune_load.argtypes = [C.c_char_p]
    L.mfista_tune_clear.restype = None
    L.mfista_tune_clear.argtypes = []
    L.mfista_tune_runs.restype = C.c_long
    L.mfista_tune_runs.argtypes = []


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _box(cl_box, n):
    if cl_box is None:
        return 0, None, C.cast(None, _fp)
    b = np.ascontiguousarray(cl_box, dtype=np.float32).ravel()
    if b.size != n:
        raise ValueError("cl_box must hold NX*NY floats")
    return 1, b, b.ctypes.data_as(_fp)


def make_opts(preset=PRESET_C24, fgp_iter=FGPITER, rel_tol=0.0, verbose=0, engine=ENGINE_PERSISTENT,
              single_trials=False, deterministic=False, autotune=False, tune_iters=30, loop=LOOP_AUTO):
    o = Opts()
    lib().mfista_default_opts(C.byref(o))
    o.preset, o.fgp_iter, o.rel_tol, o.verbose, o.engine = preset, fgp_iter, rel_tol, verbose, engine
    o.single_trials = 1 if single_trials else 0
    o.deterministic = 1 if deterministic else 0
    o.autotune = 1 if autotune else 0
    o.tune_iters = int(tune_iters)
    o.loop = int(loop)
    return o


def _info(o):
    return {"max_trials": o.max_trials, "solve_ms": o.solve_ms, "total_trials": o.total_trials,
            "kernel_ms": o.kernel_ms, "trial_rounds": o.trial_rounds, "engine_used": o.engine_used,
            "loop_used": o.loop_used}


# ---------------------------------------------------------------- solvers
def mfista_imaging_core_nufft(u_dx, v_dy, vis_r, vis_i, vis_std, NX, NY, maxiter=MAXITER, eps=EPS,
                              lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=0.0, cinit=1e4, xinit=None,
                              nonneg_flag=1, cl_box=None, preset=PRESET_C24, fgp_iter=FGPITER,
                              rel_tol=0.0, engine=ENGINE_PERSISTENT, single_trials=False, deterministic=False,
                              autotune=False, tune_iters=30, loop=LOOP_AUTO):
    """NUFFT-mode MFISTA on the GPU (c/mfista_nufft_lib.c:909).  Returns (x[NX,NY], RESULT dict, opts).
    single_trials=True evaluates one backtracking trial per round in the persistent
    engine (default: L1+TSV rounds evaluate the trials c and c*ETA together);
    deterministic=True grids without floating-point atomics (bit-identical
    repeated solves, persistent engine); autotune=True uses the autotuner's
    launch configuration for this shape (tuned first, with tune_iters
    iterations per candidate, when the table has none); loop=LOOP_AUTO runs
    the Toeplitz loop for L1 / L1+TSV square images of even size up to 512^2 (a tuned
    shape: its tuned loop), LOOP_TOEPLITZ wherever it applies, LOOP_GRIDDING
    never."""
    u, pu = _d(u_dx)
    v, pv = _d(v_dy)
    vr, pvr = _d(vis_r)
    vi, pvi = _d(vis_i)
    vs, pvs = _d(vis_std)
    M = u.size
    x0 = np.zeros(NX * NY) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(NX * NY)
    bflag, b, pb = _box(cl_box, NX * NY)
    res = RESULT()
    o = make_opts(preset, fgp_iter, rel_tol, engine=engine, single_trials=single_trials,
                  deterministic=deterministic, autotune=autotune, tune_iters=tune_iters, loop=loop)
    rc = lib().mfista_imaging_core_nufft_ex(pu, pv, pvr, pvi, pvs, M, NX, NY, maxiter, eps, lambda_l1,
                                            lambda_tv, lambda_tsv, cinit, px0,
                                            xout.ctypes.data_as(_dp), nonneg_flag, bflag, pb,
                                            C.byref(res), C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    return xout.reshape(NX, NY), res.as_dict(), _info(o)


def _pack_problems(problems):
    """mfista_nufft_problem array of a list of problem dicts; returns (array,
    buffers to keep alive, output arrays)."""
    n = len(problems)
    arr = (NufftProblem * n)()
    keep = []
    cache = {}

    def dptr(a):
        key = id(a)
        if key not in cache:
            b = np.ascontiguousarray(a, dtype=np.float64).ravel()
            cache[key] = b
            keep.append(a)  # keeps id(a) unique while cached
        return cache[key]

    outs = []
    for k, p in enumerate(problems):
        NX, NY = int(p["NX"]), int(p["NY"])
        q = arr[k]
        arrays = [dptr(p[key]) for key in ("u_dx", "v_dy", "vis_r", "vis_i", "vis_std")]
        q.u_dx, q.v_dy, q.vis_r, q.vis_i, q.vis_std = [a.ctypes.data_as(_dp) for a in arrays]
        q.M = arrays[0].size
        q.NX, q.NY = NX, NY
        q.maxiter = int(p.get("maxiter", MAXITER))
        q.eps = float(p.get("eps", EPS))
        q.lambda_l1 = float(p.get("lambda_l1", 1.0))
        q.lambda_tv = float(p.get("lambda_tv", 0.0))
        q.lambda_tsv = float(p.get("lambda_tsv", 0.0))
        q.cinit = float(p.get("cinit", 1e4))
        x0 = p.get("xinit")
        x0 = np.zeros(NX * NY) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).ravel()
        xo = np.zeros(NX * NY)
        keep += [x0, xo]
        outs.append(xo)
        q.xinit = x0.ctypes.data_as(_dp)
        q.xout = xo.ctypes.data_as(_dp)
        q.nonneg_flag = int(p.get("nonneg_flag", 1))
        bflag, b, pb = _box(p.get("cl_box"), NX * NY)
        keep.append(b)
        q.box_flag, q.cl_box = bflag, pb
        q.target_objective = float(p.get("target_objective", 0.0))
    return arr, keep, outs


def mfista_imaging_core_nufft_batch(problems, preset=PRESET_C24, fgp_iter=FGPITER, rel_tol=0.0,
                                    single_trials=False, deterministic=False, autotune=False, tune_iters=30,
                                    engine=ENGINE_PERSISTENT, loop=LOOP_AUTO):
    """Independent NUFFT solves in one persistent launch (no reference counterpart:
    a loop of c/mfista_nufft_lib.c:909 calls).  `problems` is a list of dicts with
    the keyword arguments of mfista_imaging_core_nufft (u_dx, v_dy, vis_r, vis_i,
    vis_std, NX, NY and optionally maxiter, eps, lambda_l1, lambda_tv, lambda_tsv,
    cinit, xinit, nonneg_flag, cl_box, target_objective).  Problems given the same numpy arrays for
    u_dx, v_dy, vis_r, vis_i, vis_std share one NUFFT set-up.  autotune=True:
    the autotuner's launch configuration for the batch's shape.
    engine=ENGINE_STREAM: the wide engine (every stage one launch for all
    problems; L1 and L1+TSV of one image size).  Both run Toeplitz steps when
    every problem is L1 / L1+TSV on a square image up to 512^2
    (loop=LOOP_GRIDDING: gridding steps).  Returns
    (list of x[NX,NY], list of RESULT dicts, opts dict for the whole batch)."""
    n = len(problems)
    arr, keep, outs = _pack_problems(problems)
    res = (RESULT * n)()
    o = make_opts(preset, fgp_iter, rel_tol, single_trials=single_trials, deterministic=deterministic,
                  autotune=autotune, tune_iters=tune_iters, engine=engine, loop=loop)
    rc = lib().mfista_imaging_core_nufft_batch(n, arr, res, C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    xs = [xo.reshape(int(p["NX"]), int(p["NY"])) for xo, p in zip(outs, problems)]
    return xs, [r.as_dict() for r in res], _info(o)


# ---------------------------------------------------------------- autotuner
def autotune_nufft(u_dx, v_dy, vis_r, vis_i, vis_std, NX, NY, lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=0.0,
                   cinit=1e4, xinit=None, nonneg_flag=1, preset=PRESET_C24, fgp_iter=FGPITER, iters=30,
                   deterministic=False, loop=LOOP_AUTO):
    """Times the persistent solver's launch configurations on this problem
    (`iters` iterations each) and stores the fastest for its shape.  Returns
    (choice dict, list of every candidate's dict with its us_per_iter)."""
    u, pu = _d(u_dx)
    v, pv = _d(v_dy)
    vr, pvr = _d(vis_r)
    vi, pvi = _d(vis_i)
    vs, pvs = _d(vis_std)
    x0 = np.zeros(NX * NY) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    o = make_opts(preset, fgp_iter, deterministic=deterministic, loop=loop)
    best, allc, n = Tune(), (Tune * 32)(), C.c_int(0)
    rc = lib().mfista_autotune_nufft(pu, pv, pvr, pvi, pvs, u.size, NX, NY, lambda_l1, lambda_tv, lambda_tsv,
                                     cinit, px0, nonneg_flag, C.byref(o), iters, C.byref(best), allc, 32,
                                     C.byref(n))
    if rc != 0:
        raise MfistaError(rc)
    return best.as_dict(), [allc[k].as_dict() for k in range(min(n.value, 32))]


def autotune_nufft_batch(problems, preset=PRESET_C24, fgp_iter=FGPITER, iters=30, deterministic=False):
    """autotune_nufft for a batch (the problems of mfista_imaging_core_nufft_batch)."""
    n = len(problems)
    arr, keep, outs = _pack_problems(problems)
    o = make_opts(preset, fgp_iter, deterministic=deterministic)
    best, allc, na = Tune(), (Tune * 32)(), C.c_int(0)
    rc = lib().mfista_autotune_nufft_batch(n, arr, C.byref(o), iters, C.byref(best), allc, 32, C.byref(na))
    if rc != 0:
        raise MfistaError(rc)
    return best.as_dict(), [allc[k].as_dict() for k in range(min(na.value, 32))]


def tune_table():
    """The autotuner's table: a list of entry dicts."""
    n = lib().mfista_tune_table(None, 0)
    buf = (Tune * max(n, 1))()
    n = lib().mfista_tune_table(buf, n)
    return [buf[k].as_dict() for k in range(n)]


def tune_save(path):
    rc = lib().mfista_tune_save(os.fsencode(path))
    if rc < 0:
        raise MfistaError(rc)


def tune_load(path):
    """Merges a saved table; returns the number of entries read."""
    rc = lib().mfista_tune_load(os.fsencode(path))
    if rc < 0:
        raise MfistaError(rc)
    return rc


def tune_clear():
    lib().mfista_tune_clear()


def tune_runs():
    """Candidate solves the autotuner has timed in this process."""
    return int(lib().mfista_tune_runs())


def mfista_movie_nufft(vis, u, v, noise_std, time_idx, c_box, Nx, Ny, xinit, lambda_l1, lambda_tsv, lambda_t,
                       nonneg_flag=1, cinit=1e4, maxiter=MAXITER, eps=EPS, preset=PRESET_C24, rel_tol=0.0,
                       single_trials=False, deterministic=False):
    """Coupled movie solve on the GPU with the arguments of
    matlab/MFISTA_movie_eht.m:1-4 (vis complex; time_idx an (M,) array of
    1-based frame numbers or an (M, 2) array [frame, time]; c_box None or an
    (Nx, Ny) mask; xinit None or one (Nx, Ny) image).  Returns
    (x[N_t, Nx, Ny], cost, terms, outvis, info): cost the objective at the
    start of each iteration, terms (chisq, l1, TSV, d_X), outvis the complex
    model visibilities A_t x_t in input order."""
    vis = np.asarray(vis)
    vr, pvr = _d(vis.real)
    vi, pvi = _d(vis.imag)
    uu, pu = _d(u)
    vv, pv = _d(v)
    sd, psd = _d(noise_std)
    M = uu.size
    ti = np.asarray(time_idx)
    if ti.ndim == 2:
        frames, tval = ti[:, 0], np.ascontiguousarray(ti[:, 1], dtype=np.float64)
        ptv = tval.ctypes.data_as(_dp)
    else:
        frames, tval, ptv = ti, None, None
    fi = np.ascontiguousarray(frames, dtype=np.int32)
    nt = int(fi.max())
    box = None if c_box is None else np.ascontiguousarray(c_box, dtype=np.float32).ravel()
    x0 = None if xinit is None else np.ascontiguousarray(xinit, dtype=np.float64).ravel()
    xout = np.zeros(nt * Nx * Ny)
    cost = np.zeros(int(maxiter))
    terms = np.zeros(4)
    om_r, om_i = np.zeros(M), np.zeros(M)
    iters = C.c_int(0)
    o = make_opts(preset, FGPITER, rel_tol, single_trials=single_trials, deterministic=deterministic)
    rc = lib().mfista_movie_nufft(pvr, pvi, pu, pv, psd, fi.ctypes.data_as(_ip), ptv, M,
                                  None if box is None else box.ctypes.data_as(_fp), Nx, Ny,
                                  None if x0 is None else x0.ctypes.data_as(_dp), lambda_l1, lambda_tsv,
                                  lambda_t, int(nonneg_flag), cinit, int(maxiter), eps, xout.ctypes.data_as(_dp),
                                  cost.ctypes.data_as(_dp), terms.ctypes.data_as(_dp), om_r.ctypes.data_as(_dp),
                                  om_i.ctypes.data_as(_dp), C.byref(iters), C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    info = _info(o)
    info["iterations"] = iters.value
    return (xout.reshape(nt, Nx, Ny), cost[:iters.value].copy(), terms, om_r + 1j * om_i, info)


def mfista_imaging_core_fft(u_idx, v_idx, y_r, y_i, noise_stdev, NX, NY, maxiter=MAXITER, eps=EPS,
                            lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=0.0, cinit=1e4, xinit=None,
                            nonneg_flag=1, cl_box=None, fgp_iter=FGPITER, rel_tol=0.0):
    """FFT-mode MFISTA on the GPU, FP64 (c/mfista_fft_lib.c:701).  u_idx/v_idx are
    0-based grid indices.  Returns (x[NX,NY], RESULT dict, info)."""
    ui, pui = _i(u_idx)
    vi_, pvi_ = _i(v_idx)
    yr, pyr = _d(y_r)
    yi, pyi = _d(y_i)
    ns, pns = _d(noise_stdev)
    x0 = np.zeros(NX * NY) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(NX * NY)
    bflag, b, pb = _box(cl_box, NX * NY)
    res = RESULT()
    o = make_opts(PRESET_C24, fgp_iter, rel_tol)
    rc = lib().mfista_imaging_core_fft_ex(pui, pvi_, pyr, pyi, pns, ui.size, NX, NY, maxiter, eps, lambda_l1,
                                          lambda_tv, lambda_tsv, cinit, px0, xout.ctypes.data_as(_dp),
                                          nonneg_flag, bflag, pb, C.byref(res), C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    return xout.reshape(NX, NY), res.as_dict(), _info(o)


def mfista_imaging_core_dft(y, A, NX, NY, maxiter=MAXITER, eps=EPS, lambda_l1=1.0, lambda_tv=0.0,
                            lambda_tsv=0.0, cinit=1e4, xinit=None, nonneg_flag=1, cl_box=None,
                            fgp_iter=FGPITER, rel_tol=0.0):
    """DFT-mode MFISTA on the GPU, FP64 (c/mfista_dft_lib.c:540).  A is an M x N
    array (passed column-major).  Returns (x[NX,NY], RESULT dict, info, residual)."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    M, N = A.shape
    y, py = _d(y)
    x0 = np.zeros(N) if xinit is None else np.asarray(xinit, dtype=np.float64).ravel()
    x0, px0 = _d(x0)
    xout = np.zeros(N)
    bflag, b, pb = _box(cl_box, N)
    res = RESULT()
    o = make_opts(PRESET_C24, fgp_iter, rel_tol)
    rc = lib().mfista_imaging_core_dft_ex(py, A.ctypes.data_as(_dp), M, N, NX, NY, maxiter, eps, lambda_l1,
                                          lambda_tv, lambda_tsv, cinit, px0, xout.ctypes.data_as(_dp),
                                          nonneg_flag, bflag, pb, C.byref(res), C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    resid = np.ctypeslib.as_array(res.residual, shape=(M,)).copy() if res.residual else None
    if res.residual:
        C.CDLL(None).free(C.cast(res.residual, C.c_void_p))
    return xout.reshape(NX, NY), res.as_dict(), _info(o), resid


def dft_products(A, X, y=None, adjoint=True):
    """R_k = y - A x_k and G_k = A^T R_k for the K columns of X (N x K) on the
    FP64 matrix cores.  Returns (R: M x K, G: N x K or None)."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    M, N = A.shape
    X = np.asarray(X, dtype=np.float64)
    K = X.shape[1]
    Xk = np.ascontiguousarray(X.T)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    R = np.zeros((K, M))
    G = np.zeros((K, N)) if adjoint else None
    rc = lib().mfista_dft_products(A.ctypes.data_as(_dp), M, N, K, Xk.ctypes.data_as(_dp),
                                   None if yy is None else yy.ctypes.data_as(_dp), R.ctypes.data_as(_dp),
                                   None if G is None else G.ctypes.data_as(_dp))
    if rc != 0:
        raise MfistaError(rc)
    return R.T.copy(), (None if G is None else G.T.copy())


def mfista_imaging_core_dft_batch(y, A, NX, NY, lambda_l1, lambda_tsv, cinit=1.0, maxiter=MAXITER, eps=EPS,
                                  xinit=None, nonneg_flag=1, cl_box=None, rel_tol=0.0):
    """Many DFT-mode L1(+TSV) problems sharing y and A (one per entry of
    lambda_l1 / lambda_tsv), solved in lockstep with the products on the FP64
    matrix cores.  Returns (list of x[NX,NY], list of RESULT dicts, list of
    residuals, info)."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    M, N = A.shape
    l1 = np.ascontiguousarray(np.atleast_1d(lambda_l1), dtype=np.float64)
    ltsv = np.ascontiguousarray(np.broadcast_to(lambda_tsv, l1.shape), dtype=np.float64)
    cin = np.ascontiguousarray(np.broadcast_to(cinit, l1.shape), dtype=np.float64)
    n = l1.size
    y, py = _d(y)
    x0 = None if xinit is None else np.ascontiguousarray(np.broadcast_to(np.asarray(xinit, np.float64).reshape(-1, N),
                                                                         (n, N)))
    xout = np.zeros((n, N))
    bflag, b, pb = _box(cl_box, N)
    res = (RESULT * n)()
    o = make_opts(PRESET_C24, FGPITER, rel_tol)
    rc = lib().mfista_imaging_core_dft_batch(py, A.ctypes.data_as(_dp), M, N, NX, NY, n, l1.ctypes.data_as(_dp),
                                             ltsv.ctypes.data_as(_dp), cin.ctypes.data_as(_dp), int(maxiter), eps,
                                             None if x0 is None else x0.ctypes.data_as(_dp),
                                             xout.ctypes.data_as(_dp), nonneg_flag, bflag, pb, res, C.byref(o))
    if rc != 0:
        raise MfistaError(rc)
    resid = []
    for r in res:
        resid.append(np.ctypeslib.as_array(r.residual, shape=(M,)).copy() if r.residual else None)
        if r.residual:
            C.CDLL(None).free(C.cast(r.residual, C.c_void_p))
    return [x.reshape(NX, NY) for x in xout], [r.as_dict() for r in res], resid, _info(o)


def compute_looe(A, x, yAx, NX=1, NY=1, lambda_tsv=0.0):
    """Leave-one-out error estimate of a DFT-mode solution on the GPU, FP64:
    compute_LOOE_L1 (lambda_tsv = 0, c/looe_lib.c:195) or compute_LOOE_L1_TSV
    (c/looe_lib.c:276).  A: M x N; x: the N solution values in the order of A's
    columns; yAx = y - A x.  Returns a dict: looe (-1 if the active-set Hessian
    is not positive definite), looe_m, looe_std, hessian_positive, n_active."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    M, N = A.shape
    xv, px = _d(np.asarray(x, dtype=np.float64).ravel())
    r, pr = _d(np.asarray(yAx, dtype=np.float64).ravel())
    if xv.size != N or r.size != M:
        raise ValueError("x must have N values and yAx M values")
    lm, ls = C.c_double(0.0), C.c_double(0.0)
    pd, na = C.c_int(0), C.c_int(0)
    rc = lib().mfista_looe_dft(M, N, NX, NY, lambda_tsv, A.ctypes.data_as(_dp), px, pr, C.byref(lm),
                               C.byref(ls), C.byref(pd), C.byref(na))
    if rc != 0:
        raise MfistaError(rc)
    return {"looe": lm.value if pd.value else -1.0, "looe_m": lm.value, "looe_std": ls.value,
            "hessian_positive": pd.value, "n_active": na.value}


def solve_nufft_device(t_u, t_v, t_vr, t_vi, t_sd, NX, NY, t_xinit, t_xout, maxiter=MAXITER, eps=EPS,
                       lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=0.0, cinit=1e4, nonneg_flag=1,
                       t_box=None, preset=PRESET_C24, fgp_iter=FGPITER, rel_tol=0.0, stream=None,
                       engine=ENGINE_PERSISTENT, single_trials=False, deterministic=False, autotune=False,
                       loop=LOOP_AUTO):
    """HBM-resident solve: every array argument is a device pointer (int) or a
    torch CUDA tensor (float64; cl_box float32).  The solve is queued on
    `stream` (a hipStream_t as int); by default on torch's current stream, so
    that it is ordered after the writes of the inputs.  Returns (RESULT dict,
    opts dict)."""

    def ptr(t):
        if t is None:
            return None
        return t if isinstance(t, int) else t.data_ptr()

    M = t_u.numel() if hasattr(t_u, "numel") else None
    if M is None:
        raise ValueError("pass torch tensors (M is taken from t_u)")
    if stream is None:
        import torch

        stream = torch.cuda.current_stream().cuda_stream
    res = RESULT()
    o = make_opts(preset, fgp_iter, rel_tol, engine=engine, single_trials=single_trials,
                  deterministic=deterministic, autotune=autotune, loop=loop)
    rc = lib().mfista_nufft_solve_device(ptr(t_u), ptr(t_v), ptr(t_vr), ptr(t_vi), ptr(t_sd), M, NX, NY,
                                         maxiter, eps, lambda_l1, lambda_tv, lambda_tsv, cinit,
                                         ptr(t_xinit), ptr(t_xout), nonneg_flag,
                                         0 if t_box is None else 1, ptr(t_box), C.byref(res),
                                         C.byref(o), stream)
    if rc != 0:
        raise MfistaError(rc)
    return res.as_dict(), _info(o)


def solve_nufft_device_noalloc(t_u, t_v, t_vr, t_vi, t_sd, NX, NY, t_xinit, t_xout, maxiter=MAXITER, eps=EPS,
                               lambda_l1=1.0, lambda_tv=0.0, lambda_tsv=0.0, cinit=1e4, nonneg_flag=1, t_box=None,
                               preset=PRESET_C24, fgp_iter=FGPITER, rel_tol=0.0, stream=None, deterministic=False):
    """solve_nufft_device through the allocation-free entry point: the
    library's per-device scratch array (MFISTA_NOALLOC_SCRATCH_BYTES) holds
    every workspace; no device memory is allocated.  Returns (RESULT dict,
    opts dict)."""

    def ptr(t):
        if t is None:
            return None
        return t if isinstance(t, int) else t.data_ptr()

    if stream is None:
        import torch

        stream = torch.cuda.current_stream().cuda_stream
    res = RESULT()
    o = make_opts(preset, fgp_iter, rel_tol, deterministic=deterministic)
    rc = lib().mfista_nufft_solve_device_noalloc(ptr(t_u), ptr(t_v), ptr(t_vr), ptr(t_vi), ptr(t_sd), t_u.numel(),
                                                 NX, NY, maxiter, eps, lambda_l1, lambda_tv, lambda_tsv, cinit,
                                                 ptr(t_xinit), ptr(t_xout), nonneg_flag, 0 if t_box is None else 1,
                                                 ptr(t_box), C.byref(res), C.byref(o), stream)
    if rc != 0:
        raise MfistaError(rc)
    return res.as_dict(), _info(o)


def device_allocations():
    """Device allocations the library has made since it was loaded (diagnostic)."""
    return int(lib().mfista_device_allocations())


def nufft_forward(u, v, x, preset=PRESET_C24, fp64=False):
    """Model visibilities A.x of an NX x NY image (NUFFT2d2); complex array of M."""
    u, pu = _d(u)
    v, pv = _d(v)
    x, px = _d(x)
    NX, NY = x.shape
    yr = np.zeros(u.size)
    yi = np.zeros(u.size)
    rc = lib().mfista_nufft_forward(preset, int(fp64), pu, pv, u.size, NX, NY, px,
                                    yr.ctypes.data_as(_dp), yi.ctypes.data_as(_dp))
    if rc != 0:
        raise MfistaError(rc)
    return yr + 1j * yi


def nufft_adjoint(u, v, y, NX, NY, preset=PRESET_C24, fp64=False):
    """Re(A^H y) (NUFFT2d1) as an NX x NY image."""
    u, pu = _d(u)
    v, pv = _d(v)
    yr, pyr = _d(np.real(y))
    yi, pyi = _d(np.imag(y))
    x = np.zeros((NX, NY))
    rc = lib().mfista_nufft_adjoint(preset, int(fp64), pu, pv, u.size, NX, NY, pyr, pyi,
                                    x.ctypes.data_as(_dp))
    if rc != 0:
        raise MfistaError(rc)
    return x


def x2y_nufft(u_dx, v_dy, xin):
    """c++/mfista_nufft_lib.cpp:1076 -- 12-tap model visibilities, FP64."""
    u, pu = _d(u_dx)
    v, pv = _d(v_dy)
    x, px = _d(xin)
    NX, NY = x.shape
    yr = np.zeros(u.size)
    yi = np.zeros(u.size)
    lib().x2y_nufft(pu, pv, u.size, NX, NY, px, yr.ctypes.data_as(_dp), yi.ctypes.data_as(_dp))
    return yr, yi


def y2x_nufft(u_dx, v_dy, yin_r, yin_i, NX, NY):
    """c++/mfista_nufft_lib.cpp:1151 -- 12-tap adjoint image, FP64."""
    u, pu = _d(u_dx)
    v, pv = _d(v_dy)
    yr, pyr = _d(yin_r)
    yi, pyi = _d(yin_i)
    x = np.zeros((NX, NY))
    lib().y2x_nufft(pu, pv, u.size, NX, NY, pyr, pyi, x.ctypes.data_as(_dp))
    return x


def profile_stages(reps=20):
    """Mean per-launch duration (ms) of each kernel of the last NUFFT solve,
    timed with HIP events on the library's stream."""
    ms = np.zeros(len(STAGES))
    rc = lib().mfista_profile_stages(reps, ms.ctypes.data_as(_dp), len(STAGES))
    if rc != 0:
        raise MfistaError(rc)
    return dict(zip(STAGES, ms.tolist()))


def last_steps():
    return int(lib().mfista_last_steps())


def version():
    return lib().mfista_amd_version().decode()
