"""GPU tests of the multi-right-hand-side kernel, block-diagonal batches and hipGraph replay.

Pins: the FP64 oracle per right-hand side (tolerances as in test_gpu_parity.py); a batch must
equal its matrices run one by one (bit-for-bit for the deterministic ROW kernel); every replay
of a captured graph must equal the eager result."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import device_y, rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5}


def small_mats(seed=3, count=12):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(40), np.log(3000))))
        fam = [fl.GEN_ZIPF, fl.GEN_BANDED, fl.GEN_UNIFORM][i % 3]
        pr = [(2.0, n), (9,), (4.0 / n,)][i % 3]
        mats.append(fl.gen_matrix(fam, 1000 + i, n, n, pr))
    mats.append((5, 7, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0)))  # empty block
    return mats


def spmm_oracle(m, n, cp, ri, v, X):
    k = X.shape[1]
    Y = np.zeros((m, k))
    S = np.zeros((m, k))
    for j in range(k):
        Y[:, j], S[:, j] = ob.spmv_oracle(m, n, cp, ri, v, np.ascontiguousarray(X[:, j]))
    return Y, S


def run_spmm(M, X, kind, k, accumulate=False, Y0=None):
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    dX = DeviceArray.from_host(np.ascontiguousarray(X, dt).ravel())
    dY = DeviceArray.from_host(np.ascontiguousarray(Y0, dt).ravel()) if Y0 is not None else DeviceArray.zeros(M.m * k, dt)
    M.spmm(dX, dY, k, kind, accumulate=accumulate)
    return dY.to_host().astype(np.float64).reshape(M.m, k)


@pytest.mark.parametrize("kind", [fl.VAL_F64, fl.VAL_F32], ids=["f64", "f32"])
@pytest.mark.parametrize("k", [1, 2, 8, 32, 64])
def test_spmm_row_parity(gpu, kind, k):
    rng = np.random.default_rng(k)
    for fam, m, n, pr in [(fl.GEN_ZIPF, 2000, 3000, (1.8, 500)), (fl.GEN_STENCIL27, 0, 0, (9,)),
                          (fl.GEN_BLOCKDIAG, 700, 700, (64, 0))]:
        m, n, cp, ri, v = fl.gen_matrix(fam, 11, m, n, pr)
        M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
        M.prepare_rows()
        X = rng.standard_normal((n, k))
        Y = run_spmm(M, X, kind, k)
        Yo, S = spmm_oracle(m, n, cp, ri, v, X)
        assert rel_err(Y.ravel(), Yo.ravel(), S.ravel()) <= TOL[kind]
        # accumulate mode: Y += A X
        Y0 = rng.standard_normal((m, k))
        Y2 = run_spmm(M, X, kind, k, accumulate=True, Y0=Y0)
        assert rel_err(Y2.ravel(), (Yo + Y0).ravel(), (S + np.abs(Y0)).ravel()) <= TOL[kind]


@pytest.mark.parametrize("kind", [fl.VAL_F64, fl.VAL_F32], ids=["f64", "f32"])
@pytest.mark.parametrize("k", [2, 8, 32])
def test_spmm_unaligned_blocks_match_aligned(gpu, kind, k):
    """X and Y blocks at 4- and 8-byte offsets take the narrower-lane kernels (16-byte lanes need
    16-byte alignment); every lane width sums each Y element in entry order, so the results are
    bitwise equal to the aligned run."""
    from fastload_amd._lib import check, lib
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 12, 1500, 1700, (1.8, 400))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare_rows()
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    X = np.random.default_rng(k).standard_normal((n, k)).astype(dt)
    want = run_spmm(M, X, kind, k)
    sz = np.dtype(dt).itemsize
    for off in (4, 8) if kind == fl.VAL_F32 else (8,):
        bx = DeviceArray(n * k * sz + 64, np.uint8)
        by = DeviceArray(m * k * sz + 64, np.uint8)
        check(lib.fl_memcpy_h2d(bx.ptr + off, X.ctypes.data, n * k * sz), "fl_memcpy_h2d")
        check(lib.fl_memset_device(by.ptr, 0, m * k * sz + 64), "fl_memset_device")
        check(lib.fl_spmm(M.handle, kind, k, bx.ptr + off, by.ptr + off, 0, None), "fl_spmm")
        fl.synchronize()
        got = by.to_host()[off:off + m * k * sz].view(dt).astype(np.float64).reshape(m, k)
        assert np.array_equal(got, want), (kind, k, off)


def test_spmm_invalid_k(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 1, 50, 50, (2.0, 10))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare_rows()
    with pytest.raises(fl.FastLoadError):
        run_spmm(M, np.ones((n, 3)), fl.VAL_F32, 3)


@pytest.mark.parametrize("kernel,kind", [(fl.KERNEL_PB, fl.VAL_F32), (fl.KERNEL_WIN, fl.VAL_F32),
                                         (fl.KERNEL_FAST, fl.VAL_F64), (fl.KERNEL_PB, fl.VAL_F64),
                                         (fl.KERNEL_WINR, fl.VAL_F32)],
                         ids=["pb-f32", "win-f32", "fast-f64", "pb-f64", "winr-f32"])
def test_batch_equals_individual_runs(gpu, kernel, kind):
    mats = small_mats()
    M_, N_, cp, ri, v, ro, co = fl.batch_assemble(mats)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(N_)
    B = fl.FastLoadMatrix(M_, N_, cp, ri, v, width=64)
    yb = device_y(B, x, kernel, kind, bt=128, R=4096)
    for i, (m, n, c, r, vv) in enumerate(mats):
        xi = x[co[i]:co[i + 1]]
        yo, s = ob.spmv_oracle(m, n, c, r, vv, xi)
        assert rel_err(yb[ro[i]:ro[i + 1]], yo, s) <= TOL[kind], i
        if n and len(r):
            Mi = fl.FastLoadMatrix(m, n, c, r, vv, width=64)
            yi = device_y(Mi, xi, kernel, kind, bt=128, R=4096)
            assert rel_err(yb[ro[i]:ro[i + 1]], yi, s) <= 2 * TOL[kind], i


@pytest.mark.parametrize("k", [1, 8, 32])
def test_batch_spmm_bitwise_equals_individual(gpu, k):
    """The ROW kernel is deterministic (no atomics; each row summed in entry order), so a block
    of the batch equals the same matrix run alone bit for bit."""
    mats = small_mats(seed=4, count=8)
    M_, N_, cp, ri, v, ro, co = fl.batch_assemble(mats)
    rng = np.random.default_rng(10)
    X = rng.standard_normal((N_, k))
    B = fl.FastLoadMatrix(M_, N_, cp, ri, v, width=64)
    B.prepare_rows()
    Yb = run_spmm(B, X, fl.VAL_F32, k)
    for i, (m, n, c, r, vv) in enumerate(mats):
        if not (n and len(r)):
            assert np.all(Yb[ro[i]:ro[i + 1]] == 0)
            continue
        Mi = fl.FastLoadMatrix(m, n, c, r, vv, width=64)
        Mi.prepare_rows()
        Yi = run_spmm(Mi, X[co[i]:co[i + 1]], fl.VAL_F32, k)
        assert np.array_equal(Yb[ro[i]:ro[i + 1]], Yi), i


@pytest.mark.parametrize("what", ["pb", "row"])
def test_graph_replays_match_eager(gpu, what):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 21, 20000, 20000, (1.8, 2000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    s = fl.Stream()
    rng = np.random.default_rng(2)
    if what == "pb":
        M.prepare(fl.VAL_F32)
        M.prepare_schedule(fl.KERNEL_PB, 256, 4096)
        dx = DeviceArray.from_host(rng.standard_normal(n).astype(np.float32))
        dsx = DeviceArray(n, np.float32)
        M.permute_x(dx, dsx)
        dy = DeviceArray(m, np.float32)

        def step():
            fl.zero_device(dy, s)
            M.spmv(dsx, dy, fl.KERNEL_PB, fl.VAL_F32, stream=s)
    else:
        M.prepare_rows()
        dx = DeviceArray.from_host(rng.standard_normal(n * 8).astype(np.float32))
        dy = DeviceArray(m * 8, np.float32)

        def step():
            M.spmm(dx, dy, 8, fl.VAL_F32, stream=s)
    step()  # eager: also performs first-use allocations
    s.synchronize()
    eager = dy.to_host()
    g = fl.Graph(s)
    with g:
        step()
    for rep in range(4):
        dy.copy_from(DeviceArray.from_host(np.full(dy.count, 7.0, np.float32)))  # eager work between replays
        fl.synchronize()
        g.launch()
        s.synchronize()
        got = dy.to_host()
        if what == "row":
            assert np.array_equal(got, eager), rep
        else:  # FP64 LDS sums in a different order may move the last FP32 bit
            assert np.allclose(got, eager, rtol=1e-6, atol=1e-6), rep
        step()  # eager run between replays
        s.synchronize()


@pytest.mark.parametrize("off,nbytes", [(0, 4096), (16, 1000 - 16 - 5), (3, 517), (8, 4), (0, 15), (32, 16)])
def test_zero_fill_exact_range(gpu, off, nbytes):
    """fl_zero_device clears exactly [ptr, ptr + nbytes): the 16-byte bulk, the 4-byte words and
    the byte tail, for aligned and unaligned starts."""
    from fastload_amd._lib import check, lib
    d = DeviceArray.from_host(np.full(1024, 0xAB, np.uint8))
    check(lib.fl_zero_device(d.ptr + off, nbytes, None), "fl_zero_device")
    fl.synchronize()
    h = d.to_host()
    want = np.full(1024, 0xAB, np.uint8)
    want[off:off + nbytes] = 0
    assert np.array_equal(h, want)
