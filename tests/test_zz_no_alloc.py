"""fl_spmv_assign_noalloc (the entry point that allocates nothing) through the C ABI against
fl_spmv_assign on the same prepared matrix, on two calls in a row, the second on a smaller
matrix, so a result that depended on what the previous call left in the shared static array
would show; plus the refusal of a call that needs more than the array holds."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray
from fastload_amd._lib import check

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5, fl.VAL_BF16X2: 1e-4, fl.VAL_PACKED: 2.0 ** -13}
CASES = [(fl.KERNEL_PB, fl.VAL_F32), (fl.KERNEL_PB, fl.VAL_F64), (fl.KERNEL_PB, fl.VAL_BF16X2),
         (fl.KERNEL_PB, fl.VAL_PACKED), (fl.KERNEL_FAST, fl.VAL_BF16X2), (fl.KERNEL_WIN, fl.VAL_F32),
         (fl.KERNEL_DROPIN, fl.VAL_F64), (fl.KERNEL_HYB, fl.VAL_F32), (fl.KERNEL_HYB, fl.VAL_F64),
         (fl.KERNEL_HYB, fl.VAL_PACKED)]


def _prepared(m, n, cp, ri, v, kernel, kind):
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    if kernel in (fl.KERNEL_PB, fl.KERNEL_WIN):
        M.prepare_schedule(kernel, 128, 8192)
    if kernel == fl.KERNEL_HYB:
        M.prepare_hyb(2048, 128, 8192)
    if kind != fl.VAL_F64:
        M.prepare(kind)
    return M


def _run_pair(m, n, cp, ri, v, x, kernel, kind):
    """y from fl_spmv_assign_noalloc (first, on a fresh matrix) and from fl_spmv_assign."""
    M = _prepared(m, n, cp, ri, v, kernel, kind)
    xdt = np.float64 if kind == fl.VAL_F64 else np.float32
    ydt = np.float64 if kind in (fl.VAL_F64, fl.VAL_BF16X2) else np.float32
    dx = DeviceArray.from_host(np.asarray(x, xdt))
    dsx = DeviceArray(n, xdt)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.full(m, 77.0, ydt))  # stale contents must be overwritten
    check(fl.lib.fl_spmv_assign_noalloc(M.handle, kernel, kind, dsx.ptr, dy.ptr, None), "fl_spmv_assign_noalloc")
    fl.synchronize()
    y_na = dy.to_host().astype(np.float64)
    work = DeviceArray(2 * m, np.float32) if kernel == fl.KERNEL_FAST else None
    dy2 = DeviceArray(m, ydt)
    M.spmv(dsx, dy2, kernel, kind, work=work, assign=True)
    return y_na, dy2.to_host().astype(np.float64)


@pytest.mark.parametrize("kernel,kind", CASES, ids=["pb-f32", "pb-f64", "pb-bf16x2", "pb-packed", "fast-bf16x2",
                                                    "win-f32", "dropin-f64", "hyb-f32", "hyb-f64", "hyb-packed"])
def test_noalloc_matches_assign_two_calls(gpu, kernel, kind):
    rng = np.random.default_rng(kernel * 10 + kind)
    big = fl.gen_matrix(fl.GEN_ZIPF, 31, 300000, 300000, (1.8, 4000))
    small = fl.gen_matrix(fl.GEN_ZIPF, 32, 5000, 7000, (1.8, 300))
    for (m, n, cp, ri, v) in (big, small):  # the second call's input is smaller
        x = rng.standard_normal(n)
        print(f"noalloc kernel={kernel} kind={kind} m={m} n={n} nnz={int(cp[n])}", flush=True)
        y_na, y_as = _run_pair(m, n, cp, ri, v, x, kernel, kind)
        xv = x if kind == fl.VAL_F64 else x.astype(np.float32)
        vv = v if kind in (fl.VAL_F64, fl.VAL_PACKED) else v.astype(np.float32)
        y_or, scale = ob.spmv_oracle(m, n, cp, ri, vv, xv)
        assert rel_err(y_na, y_as, scale) <= TOL[kind]
        assert rel_err(y_na, y_or, scale) <= TOL[kind]


def test_noalloc_refuses_beyond_static_array(gpu):
    """PB in FP64 on ~9.6 M entries needs a product array of ~77 MB, more than the 64 MiB
    static array: the call is refused before anything runs and y keeps its contents."""
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 33, 600000, 600000, (1.8423, 50000))
    nnz = int(cp[n])
    assert 8.5e6 < nnz < (1 << 25), nnz
    M = _prepared(m, n, cp, ri, v, fl.KERNEL_PB, fl.VAL_F64)
    dsx = DeviceArray.from_host(np.ones(n))
    dy = DeviceArray.from_host(np.full(m, 5.0))
    rc = fl.lib.fl_spmv_assign_noalloc(M.handle, fl.KERNEL_PB, fl.VAL_F64, dsx.ptr, dy.ptr, None)
    assert rc == -8
    assert b"static array" in fl.lib.fl_last_error()
    assert np.all(dy.to_host() == 5.0)
    # FP32 products (4 bytes) fit: the same matrix runs
    M.prepare(fl.VAL_F32)
    dsx32 = DeviceArray.from_host(np.ones(n, np.float32))
    dy32 = DeviceArray(m, np.float32)
    check(fl.lib.fl_spmv_assign_noalloc(M.handle, fl.KERNEL_PB, fl.VAL_F32, dsx32.ptr, dy32.ptr, None),
             "fl_spmv_assign_noalloc")
    fl.synchronize()
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), np.ones(n, np.float32))
    assert rel_err(dy32.to_host().astype(np.float64), y_or, scale) <= 1e-5
