"""CPU oracle for the masterEquation hot path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as
the thing measured or shipped.  The product path (``libtasep_hip.so``, the C
host, ``tasep_amd``) never imports this package.

Every function restates the reference (KilianPiochUBT/TASEP) in plain C; the
file:line each one follows is in ``tasep_oracle.h``.  ``ref_lib()`` returns
the reference's own ``matgen.c``/``solv.c``/``matvec.c`` compiled from source
into ``oracle/_ref`` (present only where the recipe in ``oracle/Makefile``
ran), used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libtasep_ref.so")
REF_BIN_PATH = os.path.join(HERE, "_ref", "masterEquation")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class _RunInfo(C.Structure):
    _fields_ = [("steps_done", C.c_uint64), ("steps_at_stop", C.c_uint64),
                ("early_stop", C.c_int), ("stopped", C.c_int),
                ("t", C.c_double), ("last_checksum", C.c_double)]


@dataclass
class RunInfo:
    steps_done: int
    steps_at_stop: int
    early_stop: int
    stopped: bool
    t: float
    last_checksum: float


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.tor_nnz.restype = C.c_uint64
        L.tor_nnz.argtypes = [C.c_int]
        L.tor_rowlen.restype = C.c_uint64
        L.tor_rowlen.argtypes = [C.c_int, C.c_uint64]
        L.tor_gencrs.argtypes = [C.c_int, _f64p, _u64p, _u64p, _f64p]
        L.tor_row.restype = C.c_int
        L.tor_row.argtypes = [C.c_int, C.c_uint64, _f64p, _u64p, _f64p]
        L.tor_apply_csr.argtypes = [C.c_uint64, _u64p, _u64p, _f64p, _f64p,
                                    C.c_double, _f64p, C.c_int]
        L.tor_apply_mf.argtypes = [C.c_int, _f64p, _f64p, C.c_double, _f64p]
        L.tor_step_h.restype = C.c_double
        L.tor_step_h.argtypes = [C.c_double, C.c_double, C.c_uint]
        L.tor_run_ref.restype = C.c_int
        L.tor_run_ref.argtypes = [C.c_int, _f64p, _f64p, C.c_double, C.c_double,
                                  C.c_uint64, C.c_double, C.c_int, C.c_int,
                                  C.c_int, C.POINTER(_RunInfo)]
        L.tor_tv_rates.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p,
                                   C.c_double, _i32p, _f64p]
        L.tor_run_ref_tv.restype = C.c_int
        L.tor_run_ref_tv.argtypes = [C.c_int, _f64p, C.c_double, C.c_double,
                                     C.c_uint64, _i32p, _f64p, _f64p, _f64p,
                                     _f64p, C.POINTER(_RunInfo)]
        L.tor_run_rk4.restype = C.c_int
        L.tor_run_rk4.argtypes = [C.c_int, C.c_void_p, _f64p, C.c_double,
                                  C.c_double, C.c_uint64, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.POINTER(_RunInfo)]
        L.tor_site_moments.argtypes = [C.c_int, _f64p, _f64p]
        L.tor_cov.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.tor_mass.restype = C.c_double
        L.tor_mass.argtypes = [C.c_uint64, _f64p]
        L.tor_key_sum.restype = C.c_double
        L.tor_key_sum.argtypes = [C.c_int, _f64p, C.c_uint64, C.c_int, C.c_int]
        L.tor_windows.argtypes = [C.c_int, _f64p, _f64p]
        L.tor_windows_w.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        L.tor_dense.argtypes = [C.c_int, _f64p, _f64p]
        L.tor_ref_update4.restype = C.c_double
        L.tor_ref_update4.argtypes = [C.c_uint64, _f64p, _f64p, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _info(ri: _RunInfo) -> RunInfo:
    return RunInfo(int(ri.steps_done), int(ri.steps_at_stop), int(ri.early_stop),
                   bool(ri.stopped), float(ri.t), float(ri.last_checksum))


def nnz(n: int) -> int:
    return int(lib().tor_nnz(n))


def rowlen(n: int, s: int) -> int:
    return int(lib().tor_rowlen(n, s))


def gencrs(n: int, r):
    """CSR of the generator (reference src/lib/matgen.c:7-74)."""
    dim = 1 << n
    rowPtr = np.zeros(dim + 1, np.uint64)
    colInd = np.zeros(nnz(n), np.uint64)
    val = np.zeros(nnz(n), np.float64)
    lib().tor_gencrs(n, _f64(r), rowPtr, colInd, val)
    return rowPtr, colInd, val


def row(n: int, s: int, r):
    col = np.zeros(n + 2, np.uint64)
    val = np.zeros(n + 2, np.float64)
    c = lib().tor_row(n, s, _f64(r), col, val)
    return col[:c].copy(), val[:c].copy()


def dense(n: int, r):
    dim = 1 << n
    L = np.zeros(dim * dim, np.float64)
    lib().tor_dense(n, _f64(r), L)
    return L.reshape(dim, dim)


def apply_csr(rowPtr, colInd, val, x, h, nthreads=0):
    dim = len(rowPtr) - 1
    k = np.zeros(dim, np.float64)
    lib().tor_apply_csr(dim, rowPtr, colInd, val, _f64(x), h, k, nthreads)
    return k


def apply_mf(n, r, x, h):
    k = np.zeros(1 << n, np.float64)
    lib().tor_apply_mf(n, _f64(r), _f64(x), h, k)
    return k


def step_h(T: float, t: float, steps: int) -> float:
    """h = (T - t) / (long double) steps, as masterEquation.c:586."""
    return float(lib().tor_step_h(T, t, steps))


def run_ref(n, r, x0, t0, h, steps, thr=1e-7, patience=100, use_csr=True,
            nthreads=0):
    """Reference loop with constant rates; returns (x, RunInfo)."""
    x = _f64(x0).copy()
    ri = _RunInfo()
    rc = lib().tor_run_ref(n, _f64(r), x, t0, h, steps, thr, patience,
                           1 if use_csr else 0, nthreads, C.byref(ri))
    if rc:
        raise MemoryError("oracle run_ref allocation failed")
    return x, _info(ri)


def tv_rates(n, t, tv):
    r = np.zeros(n + 1, np.float64)
    lib().tor_tv_rates(r, n + 1, _f64(tv["period"]), _f64(tv["offset"]),
                       _f64(tv["amp"]), t,
                       np.ascontiguousarray(tv["type"], np.int32), _f64(tv["c"]))
    return r


def run_ref_tv(n, x0, t0, h, steps, tv):
    x = _f64(x0).copy()
    ri = _RunInfo()
    rc = lib().tor_run_ref_tv(n, x, t0, h, steps,
                              np.ascontiguousarray(tv["type"], np.int32),
                              _f64(tv["amp"]), _f64(tv["period"]),
                              _f64(tv["offset"]), _f64(tv["c"]), C.byref(ri))
    if rc:
        raise MemoryError("oracle run_ref_tv allocation failed")
    return x, _info(ri)


def run_rk4(n, r, x0, t0, h, steps, tv=None):
    """Classical RK4 in FP64 (the build's rk4 integrator)."""
    x = _f64(x0).copy()
    ri = _RunInfo()
    keep = []

    def p(a, dt=np.float64):
        arr = np.ascontiguousarray(a, dtype=dt)
        keep.append(arr)
        return arr.ctypes.data

    if tv is None:
        args = (p(r), None, None, None, None, None)
    else:
        args = (None, p(tv["type"], np.int32), p(tv["amp"]), p(tv["period"]),
                p(tv["offset"]), p(tv["c"]))
    rc = lib().tor_run_rk4(n, args[0], x, t0, h, steps, *args[1:], C.byref(ri))
    if rc:
        raise MemoryError("oracle run_rk4 allocation failed")
    return x, _info(ri)


def site_moments(n, x):
    y = np.zeros(n, np.float64)
    lib().tor_site_moments(n, _f64(x), y)
    return y


def cov(n, x, y=None):
    """(cov, mul) in the reference's written orientation (n x n)."""
    if y is None:
        y = site_moments(n, x)
    c = np.zeros(n * n, np.float64)
    m = np.zeros(n * n, np.float64)
    lib().tor_cov(n, _f64(x), _f64(y), c, m)
    return c.reshape(n, n), m.reshape(n, n)


def key_sum(n, x, key, off, size):
    """The reference's key_fit sum over all states (masterEquation.c:32-45)."""
    return float(lib().tor_key_sum(n, _f64(x), key, off, size))


def windows(n, x, width=4):
    """W[o, p] = sum_s x_s [((s >> o) & (2^width - 1)) == p], o = 0..n-width."""
    bins = 1 << width
    W = np.zeros((n - width + 1) * bins, np.float64)
    if width == 4:
        lib().tor_windows(n, _f64(x), W)
    else:
        lib().tor_windows_w(n, width, _f64(x), W)
    return W.reshape(n - width + 1, bins)


def mass(x):
    x = _f64(x)
    return float(lib().tor_mass(len(x), x))


def uniform_x0(n):
    """Hard-coded initial state x_s = 1/2^n (masterEquation.c:402-409)."""
    return np.full(1 << n, 1.0 / float(1 << n), np.float64)


# ---------------------------------------------------------------------------
# The reference's own functions, compiled from its sources into oracle/_ref.
# ---------------------------------------------------------------------------
_ref = None


def have_ref() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("oracle/_ref not built (needs the reference sources)")
        R = C.CDLL(REF_LIB_PATH)
        R.gencrs.argtypes = [_u64p, _f64p, _u64p, _f64p, _i32p, C.c_uint64, C.c_int]
        R.rk4_step.argtypes = [C.c_int, C.c_uint64, _f64p, _f64p, _f64p, _u64p,
                               _u64p, _f64p, C.c_double]
        R.gen_meq_row.restype = C.c_uint64
        R.gen_meq_row.argtypes = [C.c_int, C.c_uint64, _f64p, _f64p, _u64p]
        R.exp_val_eval.argtypes = [C.c_uint64, _f64p, _f64p]
        R.timeVariantProbs.argtypes = [_f64p, C.c_int, _f64p, _f64p, _f64p,
                                       C.c_double, _i32p, _f64p]
        _ref = R
    return _ref
