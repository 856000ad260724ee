"""The HYB schedule on the GPU: heavy columns summed row by row from x staged in LDS, the other
columns through a PB sub-matrix whose phase 2 adds the heavy part. y is compared with the FP64
oracle to each value kind's bar (FP64 1e-12, FP32 1e-5 of sum_j |a_rj x_j|; packed: the
sub-matrix's stated rounding bound), and with the PB schedule on the same matrix; edge cases: no
heavy columns, only heavy columns, windows and slices cut by m, dense rows of heavy entries,
empty rows and columns, accumulation into y, 32-wide parents, graph replay and refusals."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5}


def _y(M, x, kind, assign=True, y0=None):
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    dx = DeviceArray.from_host(np.asarray(x, dt))
    dsx = DeviceArray(M.n, dt)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.asarray(y0 if y0 is not None else np.full(max(M.m, 1), 9.0), dt)[:max(M.m, 1)])
    M.spmv(dsx, dy, fl.KERNEL_HYB, kind, assign=assign)
    return dy.to_host()[:M.m].astype(np.float64)


def _check(m, n, cp, ri, v, x, M, kinds=(fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED), label=""):
    for kind in kinds:
        y = _y(M, x, kind)
        if kind == fl.VAL_F64:
            y_or, sc = ob.spmv_oracle(m, n, cp, ri, v, x)
            assert rel_err(y, y_or, sc) <= TOL[kind], (label, kind)
        elif kind == fl.VAL_F32:
            y_or, sc = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x.astype(np.float32))
            assert rel_err(y, y_or, sc) <= TOL[kind], (label, kind)
        else:
            y64, s64 = ob.spmv_oracle(m, n, cp, ri, v, x.astype(np.float32).astype(np.float64))
            try:
                mb = M.pack_info(fl.KERNEL_HYB)[0]
            except fl.FastLoadError:  # no light entries: everything is heavy (FP32 values)
                mb = 23
            bound = 2.0 ** -(mb + 1) * (1 + 1e-6) + 1e-6
            assert rel_err(y, y64, s64) <= bound, (label, kind, mb)


@pytest.mark.parametrize("heavy", [0, 64, 1000, 4096, 19456, 38912])
def test_hyb_power_law(gpu, heavy):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 31, 150000, 150000, (1.8, 20000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare(fl.VAL_PACKED)
    M.prepare_hyb(heavy, 128, 16384)
    H, hz = M.hyb_info()
    cnt = np.sort(np.diff(cp))
    assert H == heavy and hz == int(cnt[n - heavy:].sum())
    x = np.random.default_rng(heavy).standard_normal(n)
    kinds = (fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED) if heavy <= 19456 else (fl.VAL_F32, fl.VAL_PACKED)
    _check(m, n, cp, ri, v, x, M, kinds, f"heavy={heavy}")
    print(f"heavy {heavy}: {hz} of {int(cp[n])} entries heavy", flush=True)


@pytest.mark.parametrize("case", ["all_heavy", "cut_windows", "dense_rows", "empty_rows_cols", "stencil"])
def test_hyb_shapes(gpu, case):
    rng = np.random.default_rng(7)
    if case == "all_heavy":
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 32, 5001, 3000, (1.7, 800))
        heavy = n
    elif case == "cut_windows":  # m not a multiple of 64 or 1024
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 33, 70001, 52000, (1.9, 5000))
        heavy = 2048
    elif case == "dense_rows":  # rows holding entries of thousands of heavy columns
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF_DENSE, 34, 60000, 60000, (1.8, 5000, 0.01, 20000.0))
        heavy = 8192
    elif case == "empty_rows_cols":
        m, n = 9000, 12000
        cols = rng.integers(0, n // 2, 40000) * 2  # odd columns empty
        rows = rng.integers(0, m // 3, 40000) * 3  # two rows of three empty
        cols[:5000] = n - 2  # one long column
        A = np.zeros(n + 1, np.int64)
        order = np.lexsort((rows, cols))
        cols, rows = cols[order], rows[order]
        keep = np.ones(len(cols), bool)
        keep[1:] = (cols[1:] != cols[:-1]) | (rows[1:] != rows[:-1])
        cols, rows = cols[keep], rows[keep]
        np.add.at(A, cols + 1, 1)
        cp = np.cumsum(A).astype(np.int32)
        ri = rows.astype(np.int32)
        v = rng.standard_normal(len(ri))
        heavy = 700
    else:
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_STENCIL27, 35, 0, 0, (30,))
        heavy = min(n, 32768)  # 27000: every column
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_hyb(heavy, 64, 4096)
    M.prepare(fl.VAL_PACKED)  # after the schedule: packs the sub-matrix's PB words then
    x = rng.standard_normal(n)
    kinds = (fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED) if heavy <= 19456 else (fl.VAL_F32, fl.VAL_PACKED)
    _check(m, n, cp, ri, v, x, M, kinds, case)


@pytest.mark.parametrize("R", [16384, 8192])
def test_hyb_fold_with_partial_last_band(gpu, R):
    """m = 70,001 rows: with 16,384-row bands phase 2 reads the heavy part's y in sorted order
    (whole windows with 16-byte loads, the partial last window element by element); with 8,192
    it gathers through the row-to-position map. Both against the FP64 oracle."""
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 41, 70001, 52000, (1.9, 5000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare(fl.VAL_PACKED)
    M.prepare_hyb(4096, 128, R)
    x = np.random.default_rng(R).standard_normal(n)
    _check(m, n, cp, ri, v, x, M, (fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED), f"R={R}")


def test_hyb_matches_pb_and_accumulates(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 36, 120000, 120000, (1.8, 10000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_hyb(4096, 128, 8192)
    M.prepare_schedule(fl.KERNEL_PB, 128, 8192)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    y0 = rng.standard_normal(m)
    for kind in (fl.VAL_F64, fl.VAL_F32):
        dt = np.float64 if kind == fl.VAL_F64 else np.float32
        y_acc = _y(M, x, kind, assign=False, y0=y0)
        y_as = _y(M, x, kind, assign=True)
        y_base = y0.astype(dt).astype(np.float64)
        xv = x if kind == fl.VAL_F64 else x.astype(np.float32)
        vv = v if kind == fl.VAL_F64 else v.astype(np.float32)
        y_or, sc = ob.spmv_oracle(m, n, cp, ri, vv, xv)
        assert rel_err(y_acc - y_base, y_or, sc + np.abs(y_base)) <= TOL[kind] * 4
        assert rel_err(y_as, y_or, sc) <= TOL[kind]
        dsx = DeviceArray(n, dt)
        M.permute_x(DeviceArray.from_host(xv.astype(dt)), dsx)
        dy = DeviceArray(m, dt)
        M.spmv(dsx, dy, fl.KERNEL_PB, kind, assign=True)
        assert rel_err(y_as, dy.to_host().astype(np.float64), sc) <= 2 * TOL[kind]


def test_hyb_width32_parent_and_empty(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 37, 30000, 30000, (1.8, 3000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=32)
    M.prepare_hyb(2000, 64, 4096)
    x = np.random.default_rng(1).standard_normal(n)
    _check(m, n, cp, ri, v, x, M, (fl.VAL_F64,), "w32")
    E = fl.FastLoadMatrix(5, 7, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0), width=64)
    E.prepare_hyb(3, 64, 4096)
    assert np.array_equal(_y(E, np.ones(7), fl.VAL_F64), np.zeros(5))


def test_hyb_graph_replays(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 38, 40000, 40000, (1.8, 4000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_hyb(2048, 128, 8192)
    s = fl.Stream()
    dx = DeviceArray.from_host(np.random.default_rng(3).standard_normal(n).astype(np.float32))
    dsx = DeviceArray(n, np.float32)
    M.permute_x(dx, dsx)
    dy = DeviceArray(m, np.float32)

    def step():
        M.spmv(dsx, dy, fl.KERNEL_HYB, fl.VAL_F32, stream=s, assign=True)
    step()  # eager: first-use allocations
    s.synchronize()
    eager = dy.to_host()
    g = fl.Graph(s)
    with g:
        step()
    for rep in range(4):
        dy.copy_from(DeviceArray.from_host(np.full(m, 7.0, np.float32)))
        fl.synchronize()
        g.launch()
        s.synchronize()
        # the PB part's FP64 LDS sums may add in another order: the last FP32 bit can move
        assert np.allclose(dy.to_host(), eager, rtol=1e-6, atol=1e-6), rep
        step()
        s.synchronize()


def test_hyb_refusals(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 39, 50000, 50000, (1.8, 3000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    dsx32 = DeviceArray.from_host(np.ones(n, np.float32))
    dsx64 = DeviceArray.from_host(np.ones(n))
    y0 = np.full(m, 3.0)
    dy = DeviceArray.from_host(y0)
    with pytest.raises(fl.FastLoadError, match="HYB schedule not prepared"):
        M.spmv(dsx64, dy, fl.KERNEL_HYB, fl.VAL_F64, assign=True)
    assert np.array_equal(dy.to_host(), y0)  # refused before anything was enqueued
    for bad in ((-1, 128, 4096), (38913, 128, 4096), (100, 96, 4096), (100, 128, 512)):
        with pytest.raises(fl.FastLoadError):
            M.prepare_hyb(*bad)
    M.prepare_hyb(32768, 128, 4096)
    with pytest.raises(fl.FastLoadError, match="at most 19456 heavy columns"):
        M.spmv(dsx64, dy, fl.KERNEL_HYB, fl.VAL_F64, assign=True)
    dy32 = DeviceArray(m, np.float32)
    with pytest.raises(fl.FastLoadError, match="FP32 values not prepared"):
        M.spmv(dsx32, dy32, fl.KERNEL_HYB, fl.VAL_F32, assign=True)
    M.prepare(fl.VAL_BF16)
    with pytest.raises(fl.FastLoadError, match="supports FL_VAL_F64, FL_VAL_F32 and FL_VAL_PACKED"):
        M.spmv(dsx32, dy32, fl.KERNEL_HYB, fl.VAL_BF16, assign=True)
    M.prepare(fl.VAL_F32)
    with pytest.raises(fl.FastLoadError, match="no packed values"):
        M.spmv(dsx32, dy32, fl.KERNEL_HYB, fl.VAL_PACKED, assign=True)


def test_hyb_nonfinite_x_stays_in_its_rows(gpu):
    """An infinite x value of a heavy column reaches only the rows holding that column: inactive
    lanes of a diagonal multiply another row's entry, and their product must never be added."""
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 40, 30000, 30000, (1.8, 3000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_hyb(1024, 64, 4096)
    c = int(np.argmax(np.diff(cp)))  # the longest column: heavy
    x = np.random.default_rng(2).standard_normal(n)
    x[c] = np.inf
    hit = np.zeros(m, bool)
    hit[ri[cp[c]:cp[c + 1]]] = True
    x0 = x.copy()
    x0[c] = 0.0
    for kind in (fl.VAL_F64, fl.VAL_F32):
        y = _y(M, x, kind)
        xv = x0 if kind == fl.VAL_F64 else x0.astype(np.float32)
        vv = v if kind == fl.VAL_F64 else v.astype(np.float32)
        y_or, sc = ob.spmv_oracle(m, n, cp, ri, vv, xv)
        assert np.all(np.isfinite(y[~hit])), kind
        assert rel_err(y[~hit], y_or[~hit], sc[~hit]) <= TOL[kind], kind
        assert np.all(~np.isfinite(y[hit])), kind
