"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it was built, the
reference's own CPU twins (oracle/_ref/libref_w{32,64}.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg only. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_DIR = ROOT / "oracle" / "_ref"

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class SlideMatrix(C.Structure):
    """Field order and types of slide_matrix (src/format.h:6-20)."""
    _fields_ = [
        ("sortval", _DP), ("sortrowidx", _IP), ("tile_ptr", _IP), ("tile_len", _IP),
        ("tile_colidx", _IP), ("tile_tall", _IP), ("tile_format", _IP), ("countrow", _IP),
        ("segmentoffset", _IP), ("sortrowindex", _IP), ("tilenum", C.c_int), ("segsum", C.c_int),
    ]


def _np(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(str(ORACLE_SO))
        L.or_mmio_allinone.argtypes = [_IP, _IP, _IP, _IP, C.POINTER(_IP), C.POINTER(_IP), C.POINTER(_DP), C.c_char_p]
        L.or_mmio_allinone.restype = C.c_int
        L.or_free.argtypes = [C.c_void_p]
        L.or_csr2csc.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _i32p]
        L.or_colsort.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _i32p, _i32p, _i32p, _f64p, _f64p]
        L.or_count_levels_tiles.argtypes = [C.c_int, _i32p, C.c_int, _IP, C.POINTER(C.c_longlong)]
        L.or_count_levels_tiles.restype = C.c_int
        L.or_format_transform.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p, _i32p,
                                          _i32p, _i32p, _i32p, _i32p, _i32p, _f64p]
        L.or_format_transform.restype = C.c_int
        L.or_classify.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _IP]
        L.or_classify.restype = C.c_int
        L.or_fastload_cpu_f64.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_fastload_cpu_f32.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_spmv_csr_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_csc_f64.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_oracle.argtypes = [C.c_int, C.c_int, _i32p, _i32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, _f64p, _f64p]
        L.or_bf16_split.argtypes = [C.c_int, _f32p, _u16p, C.c_void_p]
        L.or_exclusive_scan.argtypes = [_i32p, C.c_int]
        L.or_pack_values.argtypes = [C.c_int, _f64p, C.c_int, _f32p]
        _lib = L
    return _lib


# ------------------------------------------------------------------------------- oracle API
def mmio_allinone(path: str):
    """Returns (rc, m, n, nnz, isSym, rowptr, colidx, val) -- AUX:65-232."""
    L = lib()
    m, n, nnz, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, v = _IP(), _IP(), _DP()
    rc = L.or_mmio_allinone(C.byref(m), C.byref(n), C.byref(nnz), C.byref(s), C.byref(rp), C.byref(ci),
                            C.byref(v), str(path).encode())
    if rc != 0:
        return rc, 0, 0, 0, 0, None, None, None
    out = (rc, m.value, n.value, nnz.value, s.value, _np(rp, m.value + 1, np.int32),
           _np(ci, nnz.value, np.int32), _np(v, nnz.value, np.float64))
    for p in (rp, ci, v):
        L.or_free(C.cast(p, C.c_void_p))
    return out


def csr2csc(m, n, rowptr, colidx, val):
    nnz = int(rowptr[m])
    colptr = np.zeros(n + 1, np.int32)
    rowidx = np.zeros(max(nnz, 1), np.int32)
    cval = np.zeros(max(nnz, 1), np.float64)
    npc = np.zeros(max(n, 1), np.int32)
    lib().or_csr2csc(m, n, nnz, np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(colidx, np.int32),
                     np.ascontiguousarray(val, np.float64), colptr, rowidx, cval, npc)
    return colptr, rowidx[:nnz], cval[:nnz], npc[:n]


def colsort(n, colptr, rowidx, val, x):
    nnz = int(colptr[n])
    perm = np.zeros(max(n, 1), np.int32)
    sortnnz = np.zeros(max(n, 1), np.int32)
    tmp_row = np.zeros(max(nnz, 1), np.int32)
    tmp_val = np.zeros(max(nnz, 1), np.float64)
    sortx = np.zeros(max(n, 1), np.float64)
    xx = np.ascontiguousarray(x, np.float64) if n else np.zeros(1)
    lib().or_colsort(n, np.ascontiguousarray(colptr, np.int32), np.ascontiguousarray(rowidx if nnz else np.zeros(1), np.int32),
                     np.ascontiguousarray(val if nnz else np.zeros(1), np.float64), xx, perm, sortnnz, tmp_row, tmp_val, sortx)
    return perm[:n], sortnnz[:n], tmp_row[:nnz], tmp_val[:nnz], sortx[:n]


def transform(n, nnz, sortnnz, tmp_row, tmp_val, W=32):
    """Levels, tiles and data layout (FT:8-185) plus classification (FT:188-285)."""
    L = lib()
    nl, tn = C.c_int(), C.c_longlong()
    sn = np.ascontiguousarray(sortnnz, np.int32) if n else np.zeros(1, np.int32)
    rc = L.or_count_levels_tiles(n, sn, W, C.byref(nl), C.byref(tn))
    assert rc == 0, rc
    Lv, T = nl.value, tn.value
    lvl_col = np.zeros(max(Lv, 1), np.int32); lvl_base = np.zeros(max(Lv, 1), np.int32); lvl_tall = np.zeros(max(Lv, 1), np.int32)
    tile_ptr = np.zeros(T + 1, np.int32); tile_len = np.zeros(max(T, 1), np.int32)
    tile_colidx = np.zeros(max(T, 1), np.int32); tile_depth = np.zeros(max(T, 1), np.int32)
    sortrowidx = np.zeros(max(nnz, 1), np.int32); sortval = np.zeros(max(nnz, 1), np.float64)
    tr = np.ascontiguousarray(tmp_row if nnz else np.zeros(1), np.int32)
    tv = np.ascontiguousarray(tmp_val if nnz else np.zeros(1), np.float64)
    rc = L.or_format_transform(n, nnz, sn, tr, tv, W, lvl_col, lvl_base, lvl_tall, tile_ptr, tile_len,
                               tile_colidx, tile_depth, sortrowidx, sortval)
    assert rc == 0, rc
    countrow = np.zeros(T + 1, np.int32)
    segoff = np.zeros(nnz + 1, np.int32)
    sortrowindex = np.zeros(max(nnz, 1), np.int32)
    tile_format = np.zeros(max(T, 1), np.int32)
    segsum = C.c_int()
    rc = L.or_classify(T, nnz, W, tile_ptr, tile_len, sortrowidx, countrow, segoff, sortrowindex, tile_format,
                       C.byref(segsum))
    assert rc == 0, rc
    S = segsum.value
    return dict(L=Lv, tilenum=T, lvl_col=lvl_col[:Lv], lvl_base=lvl_base[:Lv], lvl_tall=lvl_tall[:Lv],
                tile_ptr=tile_ptr, tile_len=tile_len[:T], tile_colidx=tile_colidx[:T], tile_depth=tile_depth[:T],
                sortrowidx=sortrowidx[:nnz], sortval=sortval[:nnz], countrow=countrow, segsum=S,
                segmentoffset=segoff[:S + 1], sortrowindex=sortrowindex[:nnz], tile_format=tile_format[:T])


def pipeline(m, n, colptr, rowidx, val, x, W=32):
    """ColSort + formatTransform on a CSC matrix; returns every intermediate."""
    nnz = int(colptr[n])
    perm, sortnnz, tmp_row, tmp_val, sortx = colsort(n, colptr, rowidx, val, x)
    out = transform(n, nnz, sortnnz, tmp_row, tmp_val, W)
    out.update(perm=perm, sortnnz=sortnnz, tmp_row=tmp_row, tmp_val=tmp_val, sortx=sortx, m=m, n=n, nnz=nnz, W=W)
    return out


def pack_values(v, m):
    """Values rounded (nearest even) to m mantissa bits, as the packed storage decodes them."""
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(max(v.size, 1), np.float32)
    lib().or_pack_values(int(v.size), v if v.size else np.zeros(1), int(m), out)
    return out[:v.size]


def exclusive_scan(a):
    """AUX:29-44: in-place exclusive prefix sum (returns a scanned copy)."""
    out = np.array(a, dtype=np.int32, copy=True)
    lib().or_exclusive_scan(out if out.size else np.zeros(1, np.int32), int(out.size))
    return out


def fastload_cpu_f64(t, sortx, m):
    y = np.zeros(max(m, 1), np.float64)
    lib().or_fastload_cpu_f64(t["tilenum"], t["tile_ptr"], _nz(t["tile_len"]), _nz(t["tile_colidx"]),
                              _nz(t["sortrowidx"]), _nzf(t["sortval"]), _nzf(np.asarray(sortx, np.float64)), y)
    return y[:m]


def fastload_cpu_f32(t, sortval32, sortx32, m):
    y = np.zeros(max(m, 1), np.float32)
    lib().or_fastload_cpu_f32(t["tilenum"], t["tile_ptr"], _nz(t["tile_len"]), _nz(t["tile_colidx"]),
                              _nz(t["sortrowidx"]), _nz32(sortval32), _nz32(sortx32), y)
    return y[:m]


def spmv_oracle(m, n, colptr, rowidx, val, x):
    """Compensated FP64 product and per-row scale sum_j |a_rj x_j|. val/x may be f64 or f32."""
    y = np.zeros(max(m, 1), np.float64)
    sc = np.zeros(max(m, 1), np.float64)
    val = np.ascontiguousarray(val)
    x = np.ascontiguousarray(x)
    if val.size == 0:
        val = np.zeros(1, val.dtype)
    v64 = val.ctypes.data if val.dtype == np.float64 else None
    v32 = val.ctypes.data if val.dtype == np.float32 else None
    x64 = x.ctypes.data if x.dtype == np.float64 else None
    x32 = x.ctypes.data if x.dtype == np.float32 else None
    assert (v64 or v32) and (x64 or x32)
    lib().or_spmv_oracle(m, n, np.ascontiguousarray(colptr, np.int32),
                         np.ascontiguousarray(rowidx if len(rowidx) else np.zeros(1), np.int32), v64, v32, x64, x32, y, sc)
    return y[:m], sc[:m]


def spmv_csr(m, rowptr, colidx, val, x):
    y = np.zeros(max(m, 1), np.float64)
    lib().or_spmv_csr_f64(m, np.ascontiguousarray(rowptr, np.int32), _nz(colidx), _nzf(val), _nzf(x), y)
    return y[:m]


def spmv_csc(m, n, colptr, rowidx, val, x):
    y = np.zeros(max(m, 1), np.float64)
    lib().or_spmv_csc_f64(m, n, np.ascontiguousarray(colptr, np.int32), _nz(rowidx), _nzf(val), _nzf(x), y)
    return y[:m]


def bf16_split(v32):
    v32 = np.ascontiguousarray(v32, np.float32)
    hi = np.zeros(max(v32.size, 1), np.uint16)
    lo = np.zeros(max(v32.size, 1), np.uint16)
    lib().or_bf16_split(v32.size, v32 if v32.size else np.zeros(1, np.float32), hi, lo.ctypes.data)
    return hi[:v32.size], lo[:v32.size]


def _nz(a):
    a = np.ascontiguousarray(a, np.int32)
    return a if a.size else np.zeros(1, np.int32)


def _nzf(a):
    a = np.ascontiguousarray(a, np.float64)
    return a if a.size else np.zeros(1, np.float64)


def _nz32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a if a.size else np.zeros(1, np.float32)


# ------------------------------------------------------------------- reference CPU twins
_ref = {}


def ref_available(W=32) -> bool:
    return (REF_DIR / f"libref_w{W}.so").exists()


def ref_lib(W=32):
    if W not in _ref:
        path = REF_DIR / f"libref_w{W}.so"
        if not path.exists():
            return None
        L = C.CDLL(str(path))
        L.ref_mmio_allinone.argtypes = [_IP, _IP, _IP, _IP, C.POINTER(_IP), C.POINTER(_IP), C.POINTER(_DP), C.c_char_p]
        L.ref_mmio_allinone.restype = C.c_int
        L.ref_csr2csc.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.POINTER(_IP), C.POINTER(_IP),
                                  C.POINTER(_IP), C.POINTER(_DP)]
        L.ref_col_sort.argtypes = [C.c_int, _i32p, _i32p, _i32p, _f64p, _i32p, _f64p, _i32p, _f64p, _f64p, C.c_int]
        L.ref_transform_classify.argtypes = [_i32p, _f64p, _i32p, C.c_int, C.c_int, C.c_int]
        L.ref_transform_classify.restype = C.POINTER(SlideMatrix)
        L.ref_destroy.argtypes = [C.POINTER(SlideMatrix)]
        L.ref_fastload_cpu.argtypes = [C.POINTER(SlideMatrix), C.c_int, C.c_int, C.c_int, _f64p, _f64p]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_slidesize.restype = C.c_int
        _ref[W] = L
    return _ref[W]


def ref_mmio_allinone(path: str):
    L = ref_lib(32)
    m, n, nnz, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, v = _IP(), _IP(), _DP()
    rc = L.ref_mmio_allinone(C.byref(m), C.byref(n), C.byref(nnz), C.byref(s), C.byref(rp), C.byref(ci),
                             C.byref(v), str(path).encode())
    out = (rc, m.value, n.value, nnz.value, s.value, _np(rp, m.value + 1, np.int32),
           _np(ci, nnz.value, np.int32), _np(v, nnz.value, np.float64))
    for p in (rp, ci, v):
        L.ref_free(C.cast(p, C.c_void_p))
    return out


def ref_csr2csc(m, n, rowptr, colidx, val):
    L = ref_lib(32)
    nnz = int(rowptr[m])
    ri, cp, npc, cv = _IP(), _IP(), _IP(), _DP()
    L.ref_csr2csc(nnz, m, n, _nz(colidx), np.ascontiguousarray(rowptr, np.int32), _nzf(val), C.byref(ri), C.byref(cp),
                  C.byref(npc), C.byref(cv))
    out = (_np(cp, n + 1, np.int32), _np(ri, nnz, np.int32), _np(cv, nnz, np.float64), _np(npc, n, np.int32))
    for p in (ri, cp, npc, cv):
        L.ref_free(C.cast(p, C.c_void_p))
    return out


def ref_col_sort(n, colptr, rowidx, val, x):
    L = ref_lib(32)
    nnz = int(colptr[n])
    tr = np.zeros(max(nnz, 1), np.int32); tv = np.zeros(max(nnz, 1), np.float64)
    sn = np.zeros(max(n, 1), np.int32); sx = np.zeros(max(n, 1), np.float64)
    npc = np.ascontiguousarray(np.diff(colptr), np.int32)
    L.ref_col_sort(n, npc, np.ascontiguousarray(colptr, np.int32), _nz(rowidx), _nzf(val), tr, tv, sn,
                   _nzf(x), sx, nnz)
    return sn[:n], tr[:nnz], tv[:nnz], sx[:n]


def ref_transform_classify(W, m, n, nnz, sortnnz, tmp_row, tmp_val):
    """Runs the reference's formattransation + tile_classification; returns numpy copies."""
    L = ref_lib(W)
    # formattransation reads one element past sortnnz[n-1] inside exclusive_scan (unused value);
    # pass an n+1 array so that read stays inside the allocation.
    sn = np.zeros(n + 1, np.int32)
    sn[:n] = sortnnz
    A = L.ref_transform_classify(_nz(tmp_row), _nzf(tmp_val), sn, nnz, m, n)
    a = A.contents
    T, S = a.tilenum, a.segsum
    out = dict(tilenum=T, segsum=S, tile_ptr=_np(a.tile_ptr, T + 1, np.int32), tile_len=_np(a.tile_len, T, np.int32),
               tile_colidx=_np(a.tile_colidx, T, np.int32), tile_format=_np(a.tile_format, T, np.int32),
               sortrowidx=_np(a.sortrowidx, nnz, np.int32), sortval=_np(a.sortval, nnz, np.float64),
               sortrowindex=_np(a.sortrowindex, nnz, np.int32), countrow=_np(a.countrow, T + 1, np.int32),
               segmentoffset=_np(a.segmentoffset, S + 1, np.int32))
    return A, out


def ref_fastload_cpu(A, nnz, m, n, sortx, W=32):
    L = ref_lib(W)
    y = np.zeros(max(m, 1), np.float64)
    L.ref_fastload_cpu(A, nnz, m, n, _nzf(sortx), y)
    return y[:m]


def ref_destroy(A, W=32):
    ref_lib(W).ref_destroy(A)
