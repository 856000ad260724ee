"""Helpers shared by the GPU tests: run the HIP path through the C ABI and compare with the
oracle."""
import numpy as np

import oracle_bind as ob

import fastload_amd as fl
from fastload_amd import DeviceArray


def device_y(M: "fl.FastLoadMatrix", x64, kernel, kind, bt=256, R=16384):
    """y = A x through fl_permute_x_* + fl_spmv; returns float64 host y."""
    m = M.m
    if kernel in (fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB):
        M.prepare_schedule(kernel, bt, R)
    if kind == fl.VAL_F64:
        dx = DeviceArray.from_host(np.asarray(x64, np.float64))
        dsx = DeviceArray(M.n, np.float64)
        M.permute_x(dx, dsx)
        dy = DeviceArray.zeros(m, np.float64)
        M.spmv(dsx, dy, kernel, kind)
        return dy.to_host()
    M.prepare(kind)
    dx = DeviceArray.from_host(np.asarray(x64, np.float32))
    dsx = DeviceArray(M.n, np.float32)
    M.permute_x(dx, dsx)
    if kind == fl.VAL_BF16X2:
        dy = DeviceArray.zeros(m, np.float64)
        work = DeviceArray(2 * max(m, 1), np.float32) if kernel == fl.KERNEL_FAST else None
        M.spmv(dsx, dy, kernel, kind, work=work)
        return dy.to_host()
    dy = DeviceArray.zeros(m, np.float32)
    M.spmv(dsx, dy, kernel, kind)
    return dy.to_host().astype(np.float64)


def rel_err(y, y_ref, scale):
    """max_r |y_r - ref_r| / max(scale_r, tiny): error relative to sum_j |a_rj x_j|."""
    if len(y) == 0:
        return 0.0
    den = np.maximum(scale, np.finfo(np.float64).tiny)
    err = np.abs(y - y_ref)
    err[scale == 0] = np.abs(y - y_ref)[scale == 0]  # empty rows must be exactly 0
    return float(np.max(err / den))


def export_matches_oracle(M, m, n, cp, ri, v, x, W):
    """Bit-exact comparison of every array of the device layout with the oracle."""
    got = M.export()
    t = ob.pipeline(m, n, cp, ri, v, x, W)
    keys = ["perm", "sortnnz", "tile_ptr", "tile_len", "tile_colidx", "tile_format", "sortrowidx", "sortrowindex",
            "sortval", "countrow", "segmentoffset", "lvl_col", "lvl_base", "lvl_tall"]
    bad = [k for k in keys if not np.array_equal(np.asarray(got[k]), np.asarray(t[k]))]
    assert got["tilenum"] == t["tilenum"] and got["L"] == t["L"] and got["segsum"] == t["segsum"]
    return bad, t
