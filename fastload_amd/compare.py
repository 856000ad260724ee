"""ctypes bindings of the measurement-only comparators (fastload_amd/libfastload_compare.so,
include/fastload_compare.h): the reference's naive CSC kernel (src/CSCSpMV.h) and rocSPARSE's CSR
SpMV in place of cuSPARSE (src/cusparse_cuda.h). Imported by the benchmark and the tests only; the
fastload_amd package itself never loads this library."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .device import DeviceArray

LIB_PATH = Path(__file__).resolve().parent / "libfastload_compare.so"
vp = C.c_void_p
_SIGS = {
    "flc_last_error": (C.c_char_p, []),
    "flc_csc_spmv": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_double)]),
    "flc_csc_spmv_device": (C.c_int, [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int, vp]),
    "flc_rocsparse_spmv": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                     C.POINTER(C.c_double)]),
    "flc_rocsparse_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]),
    "flc_rocsparse_run": (C.c_int, [vp, vp, vp, vp]),
    "flc_rocsparse_destroy": (C.c_int, [vp]),
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"comparator library {LIB_PATH} is missing; build it with `python fastload_amd/build.py`")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class CompareError(RuntimeError):
    def __init__(self, rc, where):
        super().__init__(f"{where} failed with status {rc}: {lib.flc_last_error().decode(errors='replace')}")
        self.rc = rc


def _check(rc, where):
    if rc < 0:
        raise CompareError(rc, where)
    return rc


def _p(a):
    return None if a is None else a.ctypes.data


def csc_spmv(m, n, colptr, rowidx, val, x, y_golden=None, filename=None):
    """src/CSCSpMV.h:29-131 -> (y, mean_ms, passed)."""
    cp = np.ascontiguousarray(colptr, np.int32)
    nnz = int(cp[n])
    ri = np.ascontiguousarray(rowidx, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(max(m, 1))
    g = None if y_golden is None else np.ascontiguousarray(y_golden, np.float64)
    t = C.c_double()
    rc = _check(lib.flc_csc_spmv(filename.encode() if filename else None, m, n, nnz, _p(cp), _p(ri), _p(v), _p(xx),
                                 _p(y), _p(g), C.byref(t)), "flc_csc_spmv")
    return y[:m], t.value, rc == 0


def rocsparse_spmv(rows, cols, rowptr, colidx, val, x, y_result=None, filename=None):
    """src/cusparse_cuda.h:27-155 with rocSPARSE -> (y, ms, passed)."""
    rp = np.ascontiguousarray(rowptr, np.int32)
    nnz = int(rp[rows])
    ci = np.ascontiguousarray(colidx, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(max(rows, 1))
    g = None if y_result is None else np.ascontiguousarray(y_result, np.float64)
    t = C.c_double()
    rc = _check(lib.flc_rocsparse_spmv(filename.encode() if filename else None, nnz, rows, cols, _p(rp), _p(ci), _p(v),
                                       _p(xx), _p(y), _p(g), C.byref(t)), "flc_rocsparse_spmv")
    return y[:rows], t.value, rc == 0


def csc_naive_device(m, n, dcolptr: DeviceArray, drowidx: DeviceArray, dval: DeviceArray, dx: DeviceArray,
                     dy: DeviceArray, stream=None):
    """y += A x with the naive CSC kernel on device arrays (fp64 when dval is float64)."""
    fp64 = dval.dtype == np.float64
    dt = np.float64 if fp64 else np.float32
    if dx.dtype != dt or dy.dtype != dt or dx.count < n or dy.count < m or dcolptr.count < n + 1:
        raise ValueError("csc_naive_device: buffer sizes or element types do not match the matrix")
    _check(lib.flc_csc_spmv_device(m, n, dcolptr.ptr, drowidx.ptr, dval.ptr, dx.ptr, dy.ptr, 1 if fp64 else 0,
                                   stream.ptr if stream is not None else None), "flc_csc_spmv_device")


class RocsparseCSR:
    """rocSPARSE adaptive CSR SpMV on resident arrays (y = A x); measurement only."""

    def __init__(self, rows, cols, drowptr: DeviceArray, dcol: DeviceArray, dval: DeviceArray, stream=None):
        self.rows, self.cols = rows, cols
        self.fp64 = dval.dtype == np.float64
        self._keep = (drowptr, dcol, dval)
        h = C.c_void_p()
        _check(lib.flc_rocsparse_create(C.byref(h), rows, cols, dcol.count, drowptr.ptr, dcol.ptr, dval.ptr,
                                        1 if self.fp64 else 0, stream.ptr if stream is not None else None),
               "flc_rocsparse_create")
        self.h = h.value

    def run(self, dx: DeviceArray, dy: DeviceArray, stream=None):
        dt = np.float64 if self.fp64 else np.float32
        if dx.dtype != dt or dy.dtype != dt or dx.count < self.cols or dy.count < self.rows:
            raise ValueError("RocsparseCSR.run: buffer sizes or element types do not match the matrix")
        _check(lib.flc_rocsparse_run(self.h, dx.ptr, dy.ptr, stream.ptr if stream is not None else None),
               "flc_rocsparse_run")

    def __del__(self):
        try:
            if self.h:
                lib.flc_rocsparse_destroy(self.h)
                self.h = None
        except Exception:
            pass
