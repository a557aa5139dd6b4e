"""ctypes binding of the single-core interval-algebra CPU baseline
(oracle/libcpustep.so, oracle/cpu_step.cpp): test infrastructure, timed by
bench.py's cpu_baseline leg and checked against the oracle."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libcpustep.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"CPU baseline not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl, P = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.POINTER
        L.cpu_last_error.restype = ctypes.c_char_p
        L.cpu_create_from_cells.restype = vp
        L.cpu_create_from_cells.argtypes = [i32, i32, i32, i32, P(ctypes.c_int32), i64]
        L.cpu_free.argtypes = [vp]
        L.cpu_nb_cells.restype = i64
        L.cpu_nb_cells.argtypes = [vp]
        L.cpu_get_cells.argtypes = [vp, P(i64)]
        L.cpu_set_values.argtypes = [vp, P(dbl)]
        L.cpu_get_values.argtypes = [vp, P(dbl)]
        L.cpu_set_dirichlet.argtypes = [vp, dbl]
        L.cpu_adapt.argtypes = [vp, dbl, dbl]
        L.cpu_upwind_step.argtypes = [vp, P(dbl), dbl]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class CpuSim:
    """One scalar field on an MR mesh, advanced by the CPU baseline."""

    def __init__(self, dim, min_level, max_level, cells, pred_s=1):
        c = np.ascontiguousarray(np.asarray(cells)[:, :4], dtype=np.int32)
        self.h = lib().cpu_create_from_cells(dim, min_level, max_level, pred_s, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             c.shape[0])
        if not self.h:
            raise RuntimeError(lib().cpu_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().cpu_free(self.h)
            self.h = None

    def nb_cells(self):
        return lib().cpu_nb_cells(self.h)

    def cells(self):
        out = np.zeros((self.nb_cells(), 5), dtype=np.int64)
        lib().cpu_get_cells(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out

    def set_values(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nb_cells()
        lib().cpu_set_values(self.h, _dp(v))

    def values(self):
        out = np.zeros(self.nb_cells())
        lib().cpu_get_values(self.h, _dp(out))
        return out

    def set_dirichlet(self, value=0.0):
        lib().cpu_set_dirichlet(self.h, float(value))

    def adapt(self, eps, reg):
        rc = lib().cpu_adapt(self.h, float(eps), float(reg))
        if rc < 0:
            raise RuntimeError(lib().cpu_last_error().decode())
        return rc

    def upwind_step(self, a, dt):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if lib().cpu_upwind_step(self.h, _dp(a), float(dt)) < 0:
            raise RuntimeError(lib().cpu_last_error().decode())
