"""fl_matrix_save / fl_matrix_load on the GPU: a device-resident matrix with its prepared
schedules written to a file and read back gives the same layout bit for bit, runs every
schedule without being prepared again, and matches the FP64 oracle; files whose indices would
leave the buffers are refused even when their hashes are right."""
import numpy as np
import pytest

import oracle_bind as ob
from flmx_util import sections, tamper
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5, fl.VAL_BF16X2: 1e-4, fl.VAL_PACKED: 2.0 ** -13}
LAYOUT_KEYS = ["perm", "sortnnz", "tile_ptr", "tile_len", "tile_colidx", "tile_format", "sortrowidx", "sortrowindex",
               "sortval", "countrow", "segmentoffset", "lvl_col", "lvl_base", "lvl_tall"]


def _y(M, x, kernel, kind):
    """y = A x on a matrix whose schedule and values are already prepared (no prepare calls)."""
    xdt = np.float64 if kind == fl.VAL_F64 else np.float32
    dx = DeviceArray.from_host(np.asarray(x, xdt))
    if kernel == -1:
        dy = DeviceArray(M.m, xdt)
        M.spmm(dx, dy, 1, kind)
        return dy.to_host().astype(np.float64)
    dsx = DeviceArray(M.n, xdt)
    M.permute_x(dx, dsx)
    dy = DeviceArray(M.m, np.float64 if kind in (fl.VAL_F64, fl.VAL_BF16X2) else np.float32)
    M.spmv(dsx, dy, kernel, kind, assign=True)
    return dy.to_host().astype(np.float64)


def _matrices():
    out = [("symgeom_w32", fl.gen_matrix(fl.GEN_SYMGEOM, 1138, 1138, 1138), 32)]
    out.append(("zipf_w64", fl.gen_matrix(fl.GEN_ZIPF, 77, 40000, 60000, (1.8, 3000)), 64))
    out.append(("stencil_w64", fl.gen_matrix(fl.GEN_STENCIL27, 5, 0, 0, (20,)), 64))
    out.append(("empty_cols", (4, 6, np.array([0, 0, 2, 2, 3, 3, 3], np.int32), np.array([1, 3, 0], np.int32),
                               np.array([1., 2, 3])), 64))
    return out


@pytest.mark.parametrize("name,mat,W", _matrices(), ids=[c[0] for c in _matrices()])
@pytest.mark.parametrize("rows_window", [False, True], ids=["col-batches", "row-window-batches"])
def test_save_load_round_trip(gpu, tmp_path, name, mat, W, rows_window):
    m, n, cp, ri, v = mat
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    runs = [(fl.KERNEL_DROPIN, fl.VAL_F64), (fl.KERNEL_DROPIN, fl.VAL_F32)]
    M.prepare(fl.VAL_F32)
    if W == 64:
        M.prepare(fl.VAL_BF16X2)
        M.prepare_rows()
        runs += [(-1, fl.VAL_F32), (-1, fl.VAL_F64), (fl.KERNEL_FAST, fl.VAL_F32)]
        if rows_window:
            M.prepare_schedule(fl.KERNEL_WINR, 64, 0)
            runs += [(fl.KERNEL_WINR, fl.VAL_F32), (fl.KERNEL_WINR, fl.VAL_F64)]
        else:
            M.prepare_schedule(fl.KERNEL_WIN, 128, 0)
            M.prepare_schedule(fl.KERNEL_PB, 128, 4096)
            M.prepare(fl.VAL_PACKED)
            runs += [(fl.KERNEL_WIN, fl.VAL_F32), (fl.KERNEL_PB, fl.VAL_F32), (fl.KERNEL_PB, fl.VAL_F64),
                     (fl.KERNEL_PB, fl.VAL_BF16X2), (fl.KERNEL_PB, fl.VAL_PACKED), (fl.KERNEL_WIN, fl.VAL_PACKED)]
    path = tmp_path / f"{name}.flmx"
    M.save(path)
    L = fl.FastLoadMatrix.load(path)
    assert (L.m, L.n, L.nnz, L.width, L.tilenum, L.levels) == (M.m, M.n, M.nnz, M.width, M.tilenum, M.levels)
    a, b = M.export(), L.export()
    for k in LAYOUT_KEYS:
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k
    x = np.random.default_rng(3).standard_normal(n)
    for kernel, kind in runs:
        y_loaded = _y(L, x, kernel, kind)
        y_orig = _y(M, x, kernel, kind)
        xv = x if kind == fl.VAL_F64 else x.astype(np.float32)
        vv = v if kind in (fl.VAL_F64, fl.VAL_PACKED) else v.astype(np.float32)
        y_or, scale = ob.spmv_oracle(m, n, cp, ri, vv, xv)
        assert rel_err(y_loaded, y_or, scale) <= TOL[kind], (kernel, kind)
        assert rel_err(y_loaded, y_orig, scale) <= TOL[kind], (kernel, kind)
        if kernel == -1:
            assert np.array_equal(y_loaded, y_orig)  # the row kernel is deterministic


@pytest.mark.parametrize("tag,index,value,fmt,what", [
    (9, 5, 10 ** 6, "i", "row index"),                 # SORTROWIDX beyond m
    (1, 0, 1, "i", "permutation"),                     # PERM with a repeated column
    (31, 7, 4096, "H", "band"),                         # PB_ROWLOCAL == R
    (30, 3, 60000, "H", "staging slots"),               # PB_LSLOT beyond the batch
    (35, 0, 10 ** 8, "i", "staging slots or runs"),     # PB_RUN_DEST beyond the product ring
    (20, 2, 65000, "H", "slots"),                       # WIN_SLOT beyond the batch's slots
    (41, 1, 10 ** 7, "i", "row order"),                 # ROW_COL beyond n
    (3, 0, 3, "i", "tile metadata"),                    # TILE_PTR differs from the recomputed layout
])
def test_load_refuses_out_of_range_indices(gpu, tmp_path, tag, index, value, fmt, what):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 78, 20000, 30000, (1.8, 2000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_rows()
    M.prepare_schedule(fl.KERNEL_WIN, 128, 0)
    M.prepare_schedule(fl.KERNEL_PB, 128, 4096)
    path = tmp_path / "good.flmx"
    M.save(path)
    raw = path.read_bytes()
    assert _loads(tmp_path, raw)
    bad = tamper(raw, tag, index, value, fmt)
    q = tmp_path / "bad.flmx"
    q.write_bytes(bytes(bad))
    with pytest.raises(fl.FastLoadError) as e:
        fl.FastLoadMatrix.load(q)
    assert e.value.rc == -8
    assert what in str(e.value), str(e.value)
    # a flipped byte without re-hashing is refused by the hash
    flip = bytearray(raw)
    flip[sections(raw)[tag][3]] ^= 1
    q.write_bytes(bytes(flip))
    with pytest.raises(fl.FastLoadError) as e:
        fl.FastLoadMatrix.load(q)
    assert "hash" in str(e.value)


def _loads(tmp_path, raw):
    q = tmp_path / "copy.flmx"
    q.write_bytes(raw)
    L = fl.FastLoadMatrix.load(q)
    return L.nnz > 0


def test_save_load_rebuilds_hyb(gpu, tmp_path):
    """A matrix saved with the HYB schedule prepared records its split and PB parameters; the
    loaded matrix runs HYB without preparing it again (FP64, FP32 and packed values) and matches
    the FP64 oracle and the original; out-of-range HYB parameters in the header are refused."""
    import struct
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 79, 50000, 60000, (1.8, 6000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_hyb(4096, 128, 8192)
    M.prepare(fl.VAL_PACKED)
    path = tmp_path / "hyb.flmx"
    M.save(path)
    L = fl.FastLoadMatrix.load(path)
    assert L.hyb_info() == M.hyb_info()
    x = np.random.default_rng(9).standard_normal(n)
    for kind in (fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED):
        dt = np.float64 if kind == fl.VAL_F64 else np.float32
        ys = []
        for A in (L, M):
            dsx = DeviceArray(n, dt)
            A.permute_x(DeviceArray.from_host(x.astype(dt)), dsx)
            dy = DeviceArray(m, dt)
            A.spmv(dsx, dy, fl.KERNEL_HYB, kind, assign=True)
            ys.append(dy.to_host().astype(np.float64))
        xv = x if kind == fl.VAL_F64 else x.astype(np.float32).astype(np.float64)
        y_or, sc = ob.spmv_oracle(m, n, cp, ri, v, xv)
        tol = 1e-12 if kind == fl.VAL_F64 else (1e-5 if kind == fl.VAL_F32 else 2.0 ** -13)
        assert rel_err(ys[0], y_or, sc) <= tol, kind
        assert rel_err(ys[0], ys[1], sc) <= tol, kind
    raw = bytearray(path.read_bytes())
    assert struct.unpack_from("<qqq", raw, 128) == (4097, 128, 8192)
    struct.pack_into("<q", raw, 128, n + 2)
    q = tmp_path / "bad_hyb.flmx"
    q.write_bytes(bytes(raw))
    with pytest.raises(fl.FastLoadError, match="HYB schedule parameters"):
        fl.FastLoadMatrix.load(q)
