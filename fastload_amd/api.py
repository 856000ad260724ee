"""Python mirror of the reference host API (same names, argument meaning and error behaviour as
src/auxiliary.h, src/csr2csc_cuda.h, src/ColSort_GPU.h, src/formatTransform_GPU.h,
src/FastLoad_GPU.h), over the C ABI of libfastload.so, plus the device-resident FastLoadMatrix.

The reference returns results through pointer/reference parameters and prints errors; here the
outputs are returned as numpy arrays and every failure raises FastLoadError.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import FastLoadError, MatrixInfo, SlideMatrix, check, i32p, f64p, lib
from .device import DeviceArray

VAL_F64, VAL_F32, VAL_BF16, VAL_BF16X2, VAL_PACKED = 0, 1, 2, 3, 4
KERNEL_DROPIN, KERNEL_FAST, KERNEL_WIN, KERNEL_PB, KERNEL_WINR, KERNEL_HYB = 0, 1, 2, 3, 4, 5

GEN_SYMGEOM, GEN_ZIPF, GEN_STENCIL27, GEN_BANDED, GEN_BLOCKDIAG, GEN_UNIFORM, GEN_ZIPF_DENSE = range(7)
VEC_NORMAL, VEC_UNIFORM, VEC_MOD10 = range(3)


def _p(a):
    return None if a is None else a.ctypes.data


def _copy_out(ptr, count, dtype):
    if count <= 0 or not ptr:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


# ------------------------------------------------------------------ reference host API
def mmio_allinone(filename: str):
    """src/auxiliary.h:65-232 -> (m, n, nnz, isSymmetric, rowptr, colidx, val)."""
    m, n, nnz, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, v = i32p(), i32p(), f64p()
    check(lib.fl_mmio_allinone(C.byref(m), C.byref(n), C.byref(nnz), C.byref(s), C.byref(rp), C.byref(ci), C.byref(v),
                               str(filename).encode()), "mmio_allinone")
    out = (m.value, n.value, nnz.value, s.value, _copy_out(rp, m.value + 1, np.int32),
           _copy_out(ci, nnz.value, np.int32), _copy_out(v, nnz.value, np.float64))
    for p in (rp, ci, v):
        lib.fl_free(C.cast(p, C.c_void_p))
    return out


def csr_to_csc(m: int, n: int, csr_val, csr_row_ptr, csr_col_ind):
    """src/csr2csc_cuda.h:51-135 -> (csc_val, csc_row_ind, csc_col_ptr, nnzpercol, time_ms)."""
    rp = np.ascontiguousarray(csr_row_ptr, np.int32)
    nnz = int(rp[m])
    ci = np.ascontiguousarray(csr_col_ind, np.int32)
    v = np.ascontiguousarray(csr_val, np.float64)
    cv = np.zeros(max(nnz, 1)); cr = np.zeros(max(nnz, 1), np.int32)
    cp = np.zeros(n + 1, np.int32); npc = np.zeros(max(n, 1), np.int32)
    t = C.c_double()
    check(lib.fl_csr_to_csc(C.byref(t), m, n, nnz, _p(v), _p(rp), _p(ci), _p(cv), _p(cr), _p(cp), _p(npc)), "csr_to_csc")
    return cv[:nnz], cr[:nnz], cp, npc[:n], t.value


def ColSort(rowA: int, colA: int, nnzPerCol, cscColPtr, cscRowIdx, cscVal, x):
    """src/ColSort_GPU.h:46-137 -> (sortrowidxTmp, sortvalTmp, sortnnzTmp, sortx, time_ms)."""
    cp = np.ascontiguousarray(cscColPtr, np.int32)
    nnz = int(cp[colA])
    npc = np.ascontiguousarray(nnzPerCol, np.int32)
    ri = np.ascontiguousarray(cscRowIdx, np.int32)
    cv = np.ascontiguousarray(cscVal, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    tr = np.zeros(max(nnz, 1), np.int32); tv = np.zeros(max(nnz, 1))
    sn = np.zeros(max(colA, 1), np.int32); sx = np.zeros(max(colA, 1))
    t = C.c_double()
    check(lib.fl_ColSort(C.byref(t), rowA, colA, nnz, _p(npc), _p(cp), _p(ri), _p(cv), _p(tr), _p(tv), _p(sn),
                         _p(xx) if colA else _p(sx), _p(sx)), "ColSort")
    return tr[:nnz], tv[:nnz], sn[:colA], sx[:colA], t.value


@dataclass
class Tiles:
    """Host copy of a slide_matrix (src/format.h:6-20)."""
    tilenum: int
    segsum: int
    tile_ptr: np.ndarray
    tile_len: np.ndarray
    tile_colidx: np.ndarray
    tile_format: np.ndarray
    sortrowidx: np.ndarray
    sortval: np.ndarray
    sortrowindex: np.ndarray
    countrow: np.ndarray
    segmentoffset: np.ndarray


def formatTransform(sortrowidx_tmp, sortval_tmp, sortnnz_tmp, nnz: int, rowA: int, colA: int):
    """src/formatTransform_GPU.h:289-596 -> (Tiles, h_count, timeFormatTran, timeFormatclas)."""
    sm = SlideMatrix()
    tr = np.ascontiguousarray(sortrowidx_tmp, np.int32)
    tv = np.ascontiguousarray(sortval_tmp, np.float64)
    sn = np.ascontiguousarray(sortnnz_tmp, np.int32)
    a, b, h = C.c_double(), C.c_double(), C.c_int()
    check(lib.fl_formatTransform(C.byref(a), C.byref(b), C.byref(sm), _p(tr) if nnz else None,
                                 _p(tv) if nnz else None, _p(sn) if colA else _p(np.zeros(1, np.int32)), nnz, rowA,
                                 colA, C.byref(h)), "formatTransform")
    try:
        T, S = sm.tilenum, sm.segsum
        tiles = Tiles(T, S, _copy_out(sm.tile_ptr, T + 1, np.int32), _copy_out(sm.tile_len, T, np.int32),
                      _copy_out(sm.tile_colidx, T, np.int32), _copy_out(sm.tile_format, T, np.int32),
                      _copy_out(sm.sortrowidx, nnz, np.int32), _copy_out(sm.sortval, nnz, np.float64),
                      _copy_out(sm.sortrowindex, nnz, np.int32), _copy_out(sm.countrow, T + 1, np.int32),
                      _copy_out(sm.segmentoffset, S + 1, np.int32))
    finally:
        lib.fl_Tile_destroy(C.byref(sm))
    return tiles, h.value, a.value, b.value


def _slide(t: Tiles):
    keep = dict(tile_ptr=np.ascontiguousarray(t.tile_ptr, np.int32), tile_len=np.ascontiguousarray(t.tile_len, np.int32),
                tile_colidx=np.ascontiguousarray(t.tile_colidx, np.int32),
                tile_format=np.ascontiguousarray(t.tile_format, np.int32),
                sortrowidx=np.ascontiguousarray(t.sortrowidx, np.int32),
                sortval=np.ascontiguousarray(t.sortval, np.float64),
                sortrowindex=np.ascontiguousarray(t.sortrowindex, np.int32))
    sm = SlideMatrix()
    for k, a in keep.items():
        setattr(sm, k, C.cast(a.ctypes.data, f64p if a.dtype == np.float64 else i32p))
    sm.tilenum = t.tilenum
    sm.segsum = t.segsum
    return sm, keep


def FastLoad_spmv(tiles: Tiles, nnz: int, colA: int, rowA: int, x, y_golden=None, filename: str | None = None):
    """src/FastLoad_GPU.h:275-409 (note the colA, rowA order) -> (y, mean_ms, passed)."""
    sm, keep = _slide(tiles)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(max(rowA, 1))
    g = None if y_golden is None else np.ascontiguousarray(y_golden, np.float64)
    t = C.c_double()
    rc = check(lib.fl_FastLoad_spmv(filename.encode() if filename else None, C.byref(sm), nnz, colA, rowA,
                                    _p(xx) if colA else _p(np.zeros(1)), _p(y), _p(g), C.byref(t)), "FastLoad_spmv")
    del keep
    return y[:rowA], t.value, rc == 0


def FastLoad_cpu(tiles: Tiles, nnz: int, rowA: int, colA: int, x, y_golden=None, threads: int = 1, y=None):
    """src/FastLoad_CPU.h:2-46 on the host -> (y, errors). x is sortx (colA entries); y (rowA
    entries, zeros unless given) accumulates the product. threads = 1 is the reference's serial
    loop (bit-identical y); 0 = every available core, > 1 that many (atomic adds, so the result
    equals the serial one up to the order of additions). With y_golden, errors counts the rows
    differing from it exactly (the reference's check) and a mismatch prints "Check NO PASS!"."""
    sm, keep = _slide(tiles)
    xx = np.ascontiguousarray(x, np.float64)
    yy = np.zeros(max(rowA, 1)) if y is None else y
    if yy.dtype != np.float64 or not yy.flags.c_contiguous or yy.size < rowA:
        raise ValueError("FastLoad_cpu: y must be a contiguous float64 array of rowA entries")
    if xx.size < colA:
        raise ValueError("FastLoad_cpu: x needs colA entries")
    g = None if y_golden is None else np.ascontiguousarray(y_golden, np.float64)
    if g is not None and g.size < rowA:
        raise ValueError("FastLoad_cpu: y_golden needs rowA entries")
    err = C.c_int()
    check(lib.fl_FastLoad_cpu(C.byref(sm), nnz, rowA, colA, _p(xx) if xx.size else _p(np.zeros(1)), _p(yy), _p(g),
                              threads, C.byref(err)), "FastLoad_cpu")
    del keep
    return yy[:rowA], err.value


def FastLoad_cpu_f32(tiles: Tiles, sortval32, nnz: int, rowA: int, colA: int, x32, threads: int = 1):
    """The FastLoad_cpu loop in FP32 (values sortval32 in tile order) -> y (float32)."""
    sm, keep = _slide(tiles)
    v = np.ascontiguousarray(sortval32, np.float32)
    xx = np.ascontiguousarray(x32, np.float32)
    if v.size < nnz or xx.size < colA:
        raise ValueError("FastLoad_cpu_f32: sortval needs nnz entries and x colA entries")
    y = np.zeros(max(rowA, 1), np.float32)
    check(lib.fl_FastLoad_cpu_f32(C.byref(sm), _p(v) if v.size else _p(np.zeros(1, np.float32)), nnz, rowA, colA,
                                  _p(xx) if xx.size else _p(np.zeros(1, np.float32)), _p(y), threads),
          "FastLoad_cpu_f32")
    del keep
    return y[:rowA]


def slide_save(path: str, tiles: Tiles, rowA: int, colA: int, nnz: int, h_count: int, perm=None):
    """Writes a slide_matrix (src/format.h:6-20) with its sizes and column order to `path`
    (fl_slide_save: the versioned FastLoad matrix file)."""
    sm, keep = _slide(tiles)
    keep["countrow"] = np.ascontiguousarray(tiles.countrow, np.int32)
    keep["segmentoffset"] = np.ascontiguousarray(tiles.segmentoffset, np.int32)
    sm.countrow = C.cast(keep["countrow"].ctypes.data, i32p)
    sm.segmentoffset = C.cast(keep["segmentoffset"].ctypes.data, i32p)
    pm = None if perm is None else np.ascontiguousarray(perm, np.int32)
    check(lib.fl_slide_save(str(path).encode(), C.byref(sm), rowA, colA, nnz, h_count, _p(pm)), "fl_slide_save")
    del keep


def slide_load(path: str):
    """Reads the slide_matrix sections of a FastLoad matrix file -> (Tiles, rowA, colA, nnz,
    h_count, perm or None)."""
    sm = SlideMatrix()
    m, n, nnz, h = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    pm = i32p()
    check(lib.fl_slide_load(str(path).encode(), C.byref(sm), C.byref(m), C.byref(n), C.byref(nnz), C.byref(h),
                            C.byref(pm)), "fl_slide_load")
    try:
        T, S, nz = sm.tilenum, sm.segsum, nnz.value
        tiles = Tiles(T, S, _copy_out(sm.tile_ptr, T + 1, np.int32), _copy_out(sm.tile_len, T, np.int32),
                      _copy_out(sm.tile_colidx, T, np.int32), _copy_out(sm.tile_format, T, np.int32),
                      _copy_out(sm.sortrowidx, nz, np.int32), _copy_out(sm.sortval, nz, np.float64),
                      _copy_out(sm.sortrowindex, nz, np.int32), _copy_out(sm.countrow, T + 1, np.int32),
                      _copy_out(sm.segmentoffset, S + 1, np.int32))
        perm = _copy_out(pm, n.value, np.int32) if pm else None
    finally:
        lib.fl_Tile_destroy(C.byref(sm))
        if pm:
            lib.fl_free(C.cast(pm, C.c_void_p))
    return tiles, m.value, n.value, nnz.value, h.value, perm


# ---------------------------------------------------------------- device-resident matrix
def _need(d: DeviceArray, count: int, dtype, what: str):
    """Refuses a device buffer too small or of the wrong element type before any kernel sees it."""
    if np.dtype(d.dtype) != np.dtype(dtype):
        raise FastLoadError(-8, f"{what}: expected {np.dtype(dtype).name}, got {np.dtype(d.dtype).name}", from_library=False)
    if d.count < count:
        raise FastLoadError(-8, f"{what}: needs {count} elements, buffer has {d.count}", from_library=False)


class FastLoadMatrix:
    """ColSort + formatTransform run on the GPU and kept resident (include/fastload.h)."""

    def __init__(self, m: int, n: int, colptr, rowidx, val, width: int = 32, stream=None):
        if isinstance(colptr, DeviceArray):
            dcp, dri, dv = colptr, rowidx, val
        else:
            dcp = DeviceArray.from_host(np.ascontiguousarray(colptr, np.int32))
            dri = DeviceArray.from_host(np.ascontiguousarray(rowidx, np.int32))
            dv = DeviceArray.from_host(np.ascontiguousarray(val, np.float64))
        nnz = dri.count if isinstance(rowidx, DeviceArray) else int(np.asarray(colptr)[n])
        h = C.c_void_p()
        check(lib.fl_matrix_create(C.byref(h), m, n, nnz, dcp.ptr, dri.ptr, dv.ptr, width,
                                   stream.ptr if stream is not None else None), "fl_matrix_create")
        self.handle = h.value
        info = MatrixInfo()
        check(lib.fl_matrix_get_info(self.handle, C.byref(info)), "fl_matrix_get_info")
        self.info = {k: getattr(info, k) for k, _ in MatrixInfo._fields_}
        self.m, self.n, self.nnz, self.width = m, n, nnz, width
        self.tilenum, self.levels = self.info["tilenum"], self.info["levels"]

    @classmethod
    def load(cls, path: str, stream=None) -> "FastLoadMatrix":
        """A matrix read back from a file written by save() (fl_matrix_load)."""
        h = C.c_void_p()
        check(lib.fl_matrix_load(C.byref(h), str(path).encode(), stream.ptr if stream is not None else None),
              "fl_matrix_load")
        M = object.__new__(cls)
        M.handle = h.value
        info = MatrixInfo()
        check(lib.fl_matrix_get_info(M.handle, C.byref(info)), "fl_matrix_get_info")
        M.info = {k: getattr(info, k) for k, _ in MatrixInfo._fields_}
        M.m, M.n, M.nnz, M.width = info.m, info.n, info.nnz, info.width
        M.tilenum, M.levels = info.tilenum, info.levels
        return M

    def save(self, path: str, stream=None):
        """Writes the layout and every prepared schedule to `path` (fl_matrix_save)."""
        check(lib.fl_matrix_save(self.handle, str(path).encode(), stream.ptr if stream is not None else None),
              "fl_matrix_save")

    def export(self) -> dict:
        i = self.info
        n, T, nz, L, S = i["n"], i["tilenum"], i["nnz"], i["levels"], i["segsum"]
        out = dict(perm=np.zeros(max(n, 1), np.int32), sortnnz=np.zeros(max(n, 1), np.int32),
                   tile_ptr=np.zeros(T + 1, np.int32), tile_len=np.zeros(max(T, 1), np.int32),
                   tile_colidx=np.zeros(max(T, 1), np.int32), tile_format=np.zeros(max(T, 1), np.int32),
                   sortrowidx=np.zeros(max(nz, 1), np.int32), sortrowindex=np.zeros(max(nz, 1), np.int32),
                   sortval=np.zeros(max(nz, 1)), countrow=np.zeros(T + 1, np.int32),
                   segmentoffset=np.zeros(S + 1, np.int32), lvl_col=np.zeros(max(L, 1), np.int32),
                   lvl_base=np.zeros(max(L, 1), np.int32), lvl_tall=np.zeros(max(L, 1), np.int32))
        order = ["perm", "sortnnz", "tile_ptr", "tile_len", "tile_colidx", "tile_format", "sortrowidx",
                 "sortrowindex", "sortval", "countrow", "segmentoffset", "lvl_col", "lvl_base", "lvl_tall"]
        check(lib.fl_matrix_export(self.handle, *[out[k].ctypes.data for k in order]), "fl_matrix_export")
        sizes = dict(perm=n, sortnnz=n, tile_len=T, tile_colidx=T, tile_format=T, sortrowidx=nz, sortrowindex=nz,
                     sortval=nz, lvl_col=L, lvl_base=L, lvl_tall=L)
        for k, s in sizes.items():
            out[k] = out[k][:s]
        out.update(tilenum=T, segsum=S, L=L)
        return out

    def perm(self) -> np.ndarray:
        """The column order (ColSort): sortx = x[perm]."""
        out = np.zeros(max(self.n, 1), np.int32)
        nulls = [None] * 13
        check(lib.fl_matrix_export(self.handle, out.ctypes.data, *nulls), "fl_matrix_export")
        return out[:self.n]

    def prepare(self, val_kind: int, stream=None):
        check(lib.fl_matrix_prepare_values(self.handle, val_kind, stream.ptr if stream is not None else None),
              "fl_matrix_prepare_values")

    def pack_info(self, kernel: int):
        """(mantissa bits, exponent-code bits, slot bits) of the schedule's packed words."""
        m, e, i = C.c_int(), C.c_int(), C.c_int()
        check(lib.fl_matrix_pack_info(self.handle, kernel, C.byref(m), C.byref(e), C.byref(i)), "fl_matrix_pack_info")
        return m.value, e.value, i.value

    def prepare_schedule(self, kernel: int, batch_tiles: int = 256, band_rows: int = 16384, stream=None):
        check(lib.fl_matrix_prepare_schedule(self.handle, kernel, batch_tiles, band_rows,
                                             stream.ptr if stream is not None else None), "fl_matrix_prepare_schedule")

    def prepare_hyb(self, heavy_cols: int | None = None, batch_tiles: int = 128, band_rows: int = 16384, stream=None):
        """The HYB schedule: the last `heavy_cols` sorted columns (default min(n, 32768)) from LDS,
        the others through a PB sub-matrix of (batch_tiles, band_rows)."""
        h = min(self.n, 32768) if heavy_cols is None else heavy_cols
        check(lib.fl_matrix_prepare_hyb(self.handle, h, batch_tiles, band_rows,
                                        stream.ptr if stream is not None else None), "fl_matrix_prepare_hyb")

    def hyb_info(self):
        """(heavy columns, heavy entries) of the prepared HYB schedule."""
        h, z = C.c_int(), C.c_longlong()
        check(lib.fl_matrix_hyb_info(self.handle, C.byref(h), C.byref(z)), "fl_matrix_hyb_info")
        return h.value, z.value

    def prepare_rows(self, stream=None):
        check(lib.fl_matrix_prepare_rows(self.handle, stream.ptr if stream is not None else None),
              "fl_matrix_prepare_rows")

    def spmm(self, dX: DeviceArray, dY: DeviceArray, k: int, val_kind=VAL_F32, accumulate=False, stream=None):
        """Y = A X (or Y += A X); X is n x k and Y is m x k, row-major, original column order."""
        if val_kind not in (VAL_F64, VAL_F32):
            raise FastLoadError(-8, "spmm: val_kind must be VAL_F64 or VAL_F32", from_library=False)
        dt = np.float64 if val_kind == VAL_F64 else np.float32
        _need(dX, self.n * k, dt, "spmm X")
        _need(dY, self.m * k, dt, "spmm Y")
        check(lib.fl_spmm(self.handle, val_kind, k, dX.ptr, dY.ptr, 1 if accumulate else 0,
                          stream.ptr if stream is not None else None), "fl_spmm")

    def permute_x(self, dx: DeviceArray, dsx: DeviceArray, stream=None):
        if dx.dtype not in (np.float64, np.float32):
            raise FastLoadError(-8, "permute_x: x must be float64 or float32", from_library=False)
        _need(dx, self.n, dx.dtype, "permute_x x")
        _need(dsx, self.n, dx.dtype, "permute_x sortx")
        fn = lib.fl_permute_x_f64 if dx.dtype == np.float64 else lib.fl_permute_x_f32
        check(fn(self.handle, dx.ptr, dsx.ptr, stream.ptr if stream is not None else None), "fl_permute_x")

    def spmv(self, dsx: DeviceArray, dy: DeviceArray, kernel=KERNEL_DROPIN, val_kind=VAL_F64, work=None, stream=None,
             assign=False):
        """y += A * sortx on the device (y = A * sortx with assign=True)."""
        if val_kind not in (VAL_F64, VAL_F32, VAL_BF16, VAL_BF16X2, VAL_PACKED):
            raise FastLoadError(-8, f"spmv: unknown val_kind {val_kind}", from_library=False)
        _need(dsx, self.n, np.float64 if val_kind == VAL_F64 else np.float32, "spmv sortx")
        _need(dy, self.m, np.float64 if val_kind in (VAL_F64, VAL_BF16X2) else np.float32, "spmv y")
        if work is not None:
            _need(work, 2 * self.m, np.float32, "spmv work")
        fn = lib.fl_spmv_assign if assign else lib.fl_spmv
        check(fn(self.handle, kernel, val_kind, dsx.ptr, dy.ptr, work.ptr if work is not None else None,
                 stream.ptr if stream is not None else None), "fl_spmv_assign" if assign else "fl_spmv")

    def close(self):
        if getattr(self, "handle", None):
            lib.fl_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------------- iterative solver
KERNEL_ROW = -1


class CG:
    """Conjugate gradients on a resident symmetric positive definite matrix (fl_cg_*): one
    prepared schedule (`kernel`, or KERNEL_ROW after prepare_rows) in FP64 or FP32, iterations
    replayed from a hipGraph in groups of `check_every`."""

    def __init__(self, M: "FastLoadMatrix", kernel: int, val_kind: int = VAL_F64):
        h = C.c_void_p()
        check(lib.fl_cg_create(C.byref(h), M.handle, kernel, val_kind), "fl_cg_create")
        self.h, self.M, self.kind = h.value, M, val_kind

    def solve(self, db: DeviceArray, dx: DeviceArray, max_iter: int = 1000, rtol: float = 1e-10,
              check_every: int = 10):
        """x (device, initial guess on entry) -> solution; returns (iterations run, ||r||/||b||)."""
        dt = np.float64 if self.kind == VAL_F64 else np.float32
        _need(db, self.M.n, dt, "cg b")
        _need(dx, self.M.n, dt, "cg x")
        it, res = C.c_int(), C.c_double()
        check(lib.fl_cg_solve(self.h, db.ptr, dx.ptr, max_iter, rtol, check_every, C.byref(it), C.byref(res)),
              "fl_cg_solve")
        return it.value, res.value

    def close(self):
        if getattr(self, "h", None):
            lib.fl_cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------- batches, graphs
def batch_assemble(mats):
    """Block-diagonal assembly of [(m, n, colptr, rowidx, val), ...] ->
    (M, N, colptr, rowidx, val, row_off, col_off) (include/fastload.h fl_batch_assemble)."""
    count = len(mats)
    keep = [(np.ascontiguousarray(c, np.int32), np.ascontiguousarray(r, np.int32), np.ascontiguousarray(v, np.float64))
            for (_, _, c, r, v) in mats]
    ms = np.array([m for (m, *_rest) in mats], np.int32)
    ns = np.array([n for (_, n, *_rest) in mats], np.int32)
    P = C.c_void_p * max(count, 1)
    cps = P(*[k[0].ctypes.data for k in keep])
    ris = P(*[k[1].ctypes.data if k[1].size else k[0].ctypes.data for k in keep])
    vvs = P(*[k[2].ctypes.data if k[2].size else k[0].ctypes.data for k in keep])
    M, N = C.c_int(), C.c_int()
    cp, ri, v = i32p(), i32p(), f64p()
    ro = np.zeros(count + 1, np.int32)
    co = np.zeros(count + 1, np.int32)
    check(lib.fl_batch_assemble(count, ms.ctypes.data, ns.ctypes.data, cps, ris, vvs, C.byref(M), C.byref(N),
                                C.byref(cp), C.byref(ri), C.byref(v), ro.ctypes.data, co.ctypes.data), "fl_batch_assemble")
    out_cp = _copy_out(cp, N.value + 1, np.int32)
    nnz = int(out_cp[-1]) if out_cp.size else 0
    out = (M.value, N.value, out_cp, _copy_out(ri, nnz, np.int32), _copy_out(v, nnz, np.float64), ro, co)
    for p in (cp, ri, v):
        lib.fl_free(C.cast(p, C.c_void_p))
    del keep
    return out


class Graph:
    """A hipGraph captured from the kernels enqueued on `stream` inside `with Graph(stream):`."""

    def __init__(self, stream):
        self.stream = stream
        self.exec = None

    def __enter__(self):
        check(lib.fl_graph_begin_capture(self.stream.ptr), "fl_graph_begin_capture")
        return self

    def __exit__(self, *exc):
        h = C.c_void_p()
        check(lib.fl_graph_end_capture(self.stream.ptr, C.byref(h)), "fl_graph_end_capture")
        self.exec = h.value
        return False

    def launch(self):
        check(lib.fl_graph_launch(self.exec, self.stream.ptr), "fl_graph_launch")

    def __del__(self):
        try:
            if self.exec:
                lib.fl_graph_destroy(self.exec)
        except Exception:
            pass


STREAM_AUTO, STREAM_NEVER, STREAM_ALWAYS = -1, 0, 1


def set_stream_policy(policy: int):
    """Cache policy of the WIN/WINR/PB/HYB read-once streams (fl_set_stream_policy): STREAM_AUTO
    (the streaming hint above the 256 MB Infinity Cache), STREAM_NEVER or STREAM_ALWAYS;
    process-wide. y does not depend on it."""
    check(lib.fl_set_stream_policy(int(policy)), "fl_set_stream_policy")


def zero_device(d: DeviceArray, stream=None):
    check(lib.fl_zero_device(d.ptr, d.nbytes, stream.ptr if stream is not None else None), "fl_zero_device")


# ------------------------------------------------------------------------- generators
def gen_matrix(family: int, seed: int, m: int, n: int, params=()):
    """Seeded synthetic CSC (include/fastload_gen.h) -> (m, n, colptr, rowidx, val)."""
    pr = np.ascontiguousarray(params, np.float64) if len(params) else None
    cp, ri, v = i32p(), i32p(), f64p()
    nz = C.c_int()
    check(lib.fl_gen_matrix(family, seed, m, n, _p(pr), 0 if pr is None else pr.size, C.byref(cp), C.byref(ri),
                            C.byref(v), C.byref(nz)), "fl_gen_matrix")
    if family == GEN_STENCIL27:
        g = int(params[0]) if len(params) else 128
        m = n = g ** 3
    out = (m, n, _copy_out(cp, n + 1, np.int32), _copy_out(ri, nz.value, np.int32), _copy_out(v, nz.value, np.float64))
    for p in (cp, ri, v):
        lib.fl_free(C.cast(p, C.c_void_p))
    return out


def gen_vector(dist: int, seed: int, count: int) -> np.ndarray:
    out = np.zeros(max(count, 1))
    check(lib.fl_gen_vector(dist, seed, count, out.ctypes.data), "fl_gen_vector")
    return out[:count]
