"""tasep_amd -- MI355X (gfx950) engine for the TASEP master-equation hot path.

Python view of the C ABI in ``include/tasep_hip.h`` (``libtasep_hip.so``).
The compute runs in hand-written HIP kernels; this module only marshals
arguments.  There is no CPU fallback: if the library or a gfx950 device is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtasep_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "tasep_hip.h")

F64, F32 = 0, 1
REF, RK4 = 0, 1
MF, CSR = 0, 1
_PREC = {"f64": F64, "f32": F32, F64: F64, F32: F32}
_INTEG = {"ref": REF, "rk4": RK4, REF: REF, RK4: RK4}
_GEN = {"mf": MF, "csr": CSR, MF: MF, CSR: CSR}

EXPORTS = [
    "tasep_ctx_create", "tasep_ctx_destroy", "tasep_set_rates",
    "tasep_set_time_variant", "tasep_set_x0_uniform", "tasep_set_x0_state",
    "tasep_set_x", "tasep_get_x", "tasep_build_csr", "tasep_get_csr",
    "tasep_csr_nnz", "tasep_gencrs", "tasep_run", "tasep_moments",
    "tasep_get_stream", "tasep_get_x_device", "tasep_strerror",
    "tasep_step_profile", "tasep_set_option", "tasep_run_resume",
    "tasep_batch_create", "tasep_batch_destroy", "tasep_batch_set_rates",
    "tasep_batch_set_time_variant", "tasep_batch_set_x0_uniform", "tasep_batch_set_x",
    "tasep_batch_get_x", "tasep_batch_run", "tasep_batch_moments",
    "tasep_batch_set_option", "tasep_batch_get_stream", "tasep_window_marginals",
    "tasep_batch_window_marginals", "tasep_batch_get_option", "tasep_window_marginals_w",
    "tasep_set_x0_product", "tasep_get_x_range", "tasep_moments_noalloc",
]
OPT_TILED, OPT_XCD_ORDER, OPT_RK4_STAGED, OPT_TILE_BLOCKS, OPT_LDS, OPT_LDS_THREADS = 1, 2, 3, 4, 5, 6
OPT_GRAPH, OPT_STORE = 7, 8
BATCH_OPT_GRAPH_STEPS, BATCH_OPT_LDS, BATCH_OPT_CONCURRENT, BATCH_OPT_SCHEDULE = 1, 2, 3, 4
BATCH_OPT_RANGE_TILES, BATCH_OPT_STORE = 5, 6


class TasepError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tasep error {code}: {msg}")
        self.code = code


class _KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("ms", C.c_double),
                ("bytes_per_state", C.c_double)]


class _RunInfo(C.Structure):
    _fields_ = [("steps_done", C.c_uint64), ("steps_at_stop", C.c_uint64),
                ("early_stop", C.c_int32), ("stopped", C.c_int32),
                ("t", C.c_double), ("last_checksum", C.c_double),
                ("mass", C.c_double)]


@dataclass
class RunInfo:
    steps_done: int
    steps_at_stop: int
    early_stop: int
    stopped: bool
    t: float
    last_checksum: float
    mass: float


_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_lib = None


def load_library():
    """Load libtasep_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C tasep_amd`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.tasep_strerror.restype = C.c_char_p
    L.tasep_strerror.argtypes = [C.c_int]
    L.tasep_csr_nnz.restype = C.c_uint64
    L.tasep_csr_nnz.argtypes = [C.c_int]
    L.tasep_ctx_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int]
    L.tasep_ctx_destroy.argtypes = [vp]
    L.tasep_set_rates.argtypes = [vp, _f64p]
    L.tasep_set_time_variant.argtypes = [vp, _i32p, _f64p, _f64p, _f64p, _f64p]
    L.tasep_set_x0_uniform.argtypes = [vp]
    L.tasep_set_x0_state.argtypes = [vp, C.c_uint64]
    L.tasep_set_x.argtypes = [vp, _f64p]
    L.tasep_get_x.argtypes = [vp, _f64p]
    L.tasep_set_x0_product.argtypes = [vp, _f64p]
    L.tasep_get_x_range.argtypes = [vp, C.c_uint64, C.c_uint64, _f64p]
    L.tasep_build_csr.argtypes = [vp]
    L.tasep_get_csr.argtypes = [vp, _u64p, _u64p, _f64p]
    L.tasep_gencrs.argtypes = [_u64p, _f64p, _u64p, _f64p, C.c_uint64, C.c_int]
    L.tasep_run.argtypes = [vp, C.c_double, C.c_double, C.c_uint64, C.c_double,
                            C.c_int, C.POINTER(_RunInfo)]
    L.tasep_run_resume.argtypes = [vp, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                   C.c_int, C.c_int, C.POINTER(_RunInfo)]
    L.tasep_moments.argtypes = [vp, vp, vp, vp, vp]
    L.tasep_moments_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.tasep_step_profile.argtypes = [vp, C.c_double, C.c_double, C.c_int,
                                     C.POINTER(_KernelTime), C.c_int,
                                     C.POINTER(C.c_int)]
    L.tasep_set_option.argtypes = [vp, C.c_int, C.c_int64]
    L.tasep_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.tasep_get_x_device.argtypes = [vp, C.POINTER(vp)]
    L.tasep_batch_create.argtypes = [C.POINTER(vp), C.c_int, _i32p, C.c_int]
    L.tasep_batch_destroy.argtypes = [vp]
    L.tasep_batch_set_rates.argtypes = [vp, C.c_int, _f64p]
    L.tasep_batch_set_time_variant.argtypes = [vp, C.c_int, _i32p, _f64p, _f64p, _f64p, _f64p]
    L.tasep_batch_set_x0_uniform.argtypes = [vp]
    L.tasep_batch_set_x.argtypes = [vp, C.c_int, _f64p]
    L.tasep_batch_get_x.argtypes = [vp, C.c_int, _f64p]
    L.tasep_batch_run.argtypes = [vp, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                  C.c_int, C.POINTER(_RunInfo)]
    L.tasep_batch_moments.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    L.tasep_batch_set_option.argtypes = [vp, C.c_int, C.c_int64]
    L.tasep_batch_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.tasep_batch_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.tasep_window_marginals.argtypes = [vp, _f64p]
    L.tasep_window_marginals_w.argtypes = [vp, C.c_int, _f64p]
    L.tasep_batch_window_marginals.argtypes = [vp, C.c_int, _f64p]
    for name in EXPORTS:
        if name not in ("tasep_strerror", "tasep_csr_nnz"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise TasepError(rc, load_library().tasep_strerror(rc).decode())


def csr_nnz(n: int) -> int:
    return int(load_library().tasep_csr_nnz(n))


def gencrs(n: int, rates):
    """Device-built CSR, same arrays as the reference's gencrs."""
    dim = 1 << n
    nnz = csr_nnz(n)
    rowPtr = np.zeros(dim + 1, np.uint64)
    colInd = np.zeros(nnz, np.uint64)
    val = np.zeros(nnz, np.float64)
    _check(load_library().tasep_gencrs(rowPtr, val, colInd,
                                       np.ascontiguousarray(rates, np.float64), dim, n))
    return rowPtr, colInd, val


class Engine:
    """One master-equation instance resident in HBM."""

    def __init__(self, n: int, precision="f64", integrator="ref", generator="mf"):
        L = load_library()
        self.n = n
        self.dim = 1 << n
        self.precision = _PREC[precision]
        self.integrator = _INTEG[integrator]
        self.generator = _GEN[generator]
        h = C.c_void_p()
        _check(L.tasep_ctx_create(C.byref(h), n, self.precision, self.integrator,
                                  self.generator))
        self._h = h

    # -- lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().tasep_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs
    def set_rates(self, r):
        r = np.ascontiguousarray(r, np.float64)
        if r.shape != (self.n + 1,):
            raise ValueError(f"need {self.n + 1} rates")
        _check(load_library().tasep_set_rates(self._h, r))

    def set_time_variant(self, tv):
        f = lambda k: np.ascontiguousarray(tv[k], np.float64)  # noqa: E731
        _check(load_library().tasep_set_time_variant(
            self._h, np.ascontiguousarray(tv["type"], np.int32), f("amp"),
            f("period"), f("offset"), f("c")))

    def set_x0_uniform(self):
        _check(load_library().tasep_set_x0_uniform(self._h))

    def set_x0_state(self, s: int):
        _check(load_library().tasep_set_x0_state(self._h, s))

    def set_x0_product(self, p):
        """x_s = prod_k (b_k(s) ? p[k] : 1 - p[k]); p[k] for bit k (site n-k)."""
        p = np.ascontiguousarray(p, np.float64)
        if p.shape != (self.n,):
            raise ValueError(f"need {self.n} site densities")
        _check(load_library().tasep_set_x0_product(self._h, p))

    def set_x(self, x):
        x = np.ascontiguousarray(x, np.float64)
        if x.shape != (self.dim,):
            raise ValueError("x must have 2^n entries")
        _check(load_library().tasep_set_x(self._h, x))

    def get_x(self):
        x = np.empty(self.dim, np.float64)
        _check(load_library().tasep_get_x(self._h, x))
        return x

    def get_x_range(self, offset: int, count: int):
        x = np.empty(count, np.float64)
        _check(load_library().tasep_get_x_range(self._h, offset, count, x))
        return x

    def set_option(self, key: int, value: int):
        _check(load_library().tasep_set_option(self._h, key, value))

    # -- generator
    def build_csr(self):
        _check(load_library().tasep_build_csr(self._h))

    def get_csr(self):
        nnz = csr_nnz(self.n)
        rp = np.zeros(self.dim + 1, np.uint64)
        ci = np.zeros(nnz, np.uint64)
        va = np.zeros(nnz, np.float64)
        _check(load_library().tasep_get_csr(self._h, rp, ci, va))
        return rp, ci, va

    # -- loop
    def run(self, t0, h, steps, thr=1e-7, patience=100, resume=False) -> RunInfo:
        ri = _RunInfo()
        _check(load_library().tasep_run_resume(self._h, t0, h, steps, thr, patience,
                                               1 if resume else 0, C.byref(ri)))
        return RunInfo(int(ri.steps_done), int(ri.steps_at_stop), int(ri.early_stop),
                       bool(ri.stopped), float(ri.t), float(ri.last_checksum),
                       float(ri.mass))

    def moments(self):
        """(y[n] by bit, cov[n,n] by bit, mul[n,n] by bit, mass)."""
        n = self.n
        y = np.zeros(n)
        cov = np.zeros(n * n)
        mul = np.zeros(n * n)
        m = np.zeros(1)
        _check(load_library().tasep_moments(self._h, y.ctypes.data, cov.ctypes.data,
                                            mul.ctypes.data, m.ctypes.data))
        return y, cov.reshape(n, n), mul.reshape(n, n), float(m[0])

    def window_marginals(self, width=4):
        """W[o, p] = sum_s x_s [((s >> o) & (2^width - 1)) == p],
        o = 0..n-width, width 4 or 6."""
        bins = 1 << width
        W = np.zeros((self.n - width + 1) * bins)
        if width == 4:
            _check(load_library().tasep_window_marginals(self._h, W))
        else:
            _check(load_library().tasep_window_marginals_w(self._h, width, W))
        return W.reshape(self.n - width + 1, bins)

    def step_profile(self, t0, h, reps=5):
        """[(name, mean ms, algorithmic bytes/state)] for each launch of a step."""
        arr = (_KernelTime * 8)()
        cnt = C.c_int()
        _check(load_library().tasep_step_profile(self._h, t0, h, reps, arr, 8,
                                                 C.byref(cnt)))
        return [(arr[i].name.decode(), arr[i].ms, arr[i].bytes_per_state)
                for i in range(cnt.value)]

    # -- device handles
    @property
    def stream(self) -> int:
        p = C.c_void_p()
        _check(load_library().tasep_get_stream(self._h, C.byref(p)))
        return int(p.value or 0)

    @property
    def x_device_ptr(self) -> int:
        p = C.c_void_p()
        _check(load_library().tasep_get_x_device(self._h, C.byref(p)))
        return int(p.value or 0)


def _info(ri) -> RunInfo:
    return RunInfo(int(ri.steps_done), int(ri.steps_at_stop), int(ri.early_stop),
                   bool(ri.stopped), float(ri.t), float(ri.last_checksum), float(ri.mass))


class Batch:
    """Many independent instances (own n and rates) driven together."""

    def __init__(self, ns, precision="f32"):
        L = load_library()
        self.ns = [int(v) for v in ns]
        self.count = len(self.ns)
        self.precision = _PREC[precision]
        h = C.c_void_p()
        _check(L.tasep_batch_create(C.byref(h), self.count,
                                    np.ascontiguousarray(self.ns, np.int32), self.precision))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            load_library().tasep_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rates(self, i, r):
        r = np.ascontiguousarray(r, np.float64)
        if 0 <= i < self.count and r.shape != (self.ns[i] + 1,):
            raise ValueError(f"instance {i} needs {self.ns[i] + 1} rates")
        _check(load_library().tasep_batch_set_rates(self._h, i, r))

    def set_time_variant(self, i, tv):
        f = lambda k: np.ascontiguousarray(tv[k], np.float64)  # noqa: E731
        _check(load_library().tasep_batch_set_time_variant(
            self._h, i, np.ascontiguousarray(tv["type"], np.int32), f("amp"), f("period"),
            f("offset"), f("c")))

    def set_x0_uniform(self):
        _check(load_library().tasep_batch_set_x0_uniform(self._h))

    def set_x(self, i, x):
        x = np.ascontiguousarray(x, np.float64)
        if x.shape != (1 << self.ns[i],):
            raise ValueError("x must have 2^n entries")
        _check(load_library().tasep_batch_set_x(self._h, i, x))

    def get_x(self, i):
        x = np.empty(1 << self.ns[i], np.float64)
        _check(load_library().tasep_batch_get_x(self._h, i, x))
        return x

    def set_option(self, key: int, value: int):
        _check(load_library().tasep_batch_set_option(self._h, key, value))

    def get_option(self, key: int) -> int:
        v = C.c_int64()
        _check(load_library().tasep_batch_get_option(self._h, key, C.byref(v)))
        return int(v.value)

    def run(self, t0, h, steps, thr=1e-7, patience=-1):
        arr = (_RunInfo * self.count)()
        _check(load_library().tasep_batch_run(self._h, t0, h, steps, thr, patience, arr))
        return [_info(arr[i]) for i in range(self.count)]

    def moments(self, i):
        n = self.ns[i]
        y = np.zeros(n)
        cov = np.zeros(n * n)
        mul = np.zeros(n * n)
        m = np.zeros(1)
        _check(load_library().tasep_batch_moments(self._h, i, y.ctypes.data, cov.ctypes.data,
                                                  mul.ctypes.data, m.ctypes.data))
        return y, cov.reshape(n, n), mul.reshape(n, n), float(m[0])

    def window_marginals(self, i):
        W = np.zeros(max(self.ns[i] - 3, 0) * 16)
        _check(load_library().tasep_batch_window_marginals(self._h, i, W))
        return W.reshape(-1, 16)

    @property
    def stream(self) -> int:
        p = C.c_void_p()
        _check(load_library().tasep_batch_get_stream(self._h, C.byref(p)))
        return int(p.value or 0)
