"""GPU parity at the benchmark's full sizes: the power-law matrix (4,194,304^2, 66.6M nnz, seed
42) and the 128^3 27-point stencil (2,097,152^2, 55.7M nnz). The oracle runs the whole
preprocessing and the compensated FP64 product at these sizes in seconds, so the layout is
checked bit for bit and y against the oracle directly; two size-independent properties are
checked as well: the checksum 1'(Ax) = (1'A)x and linearity in x."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import device_y, export_matches_oracle, rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray
from fastload_amd import selector as sel

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5, fl.VAL_BF16X2: 1e-4}


@pytest.fixture(scope="module")
def powerlaw():
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 42, 4194304, 4194304, (1.8423, 50000))
    assert int(cp[n]) == 66577983
    return m, n, cp, ri, v, fl.gen_vector(fl.VEC_NORMAL, 43, n)


@pytest.fixture(scope="module")
def stencil():
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_STENCIL27, 128, 0, 0, (128,))
    assert int(cp[n]) == 55742968
    return m, n, cp, ri, v, fl.gen_vector(fl.VEC_NORMAL, 129, n)


@pytest.mark.parametrize("W", [32, 64])
@pytest.mark.parametrize("which", ["powerlaw", "stencil"])
def test_full_size_layout_bit_exact(gpu, request, which, W):
    m, n, cp, ri, v, x = request.getfixturevalue(which)
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    bad, t = export_matches_oracle(M, m, n, cp, ri, v, x, W)
    assert not bad, f"{which} W={W}: arrays differ: {bad}"


SCHEDULES = [(32, fl.KERNEL_DROPIN, fl.VAL_F64), (32, fl.KERNEL_DROPIN, fl.VAL_F32), (64, fl.KERNEL_FAST, fl.VAL_F32),
             (64, fl.KERNEL_WIN, fl.VAL_F32), (64, fl.KERNEL_PB, fl.VAL_F32), (64, fl.KERNEL_PB, fl.VAL_BF16X2),
             (64, fl.KERNEL_PB, fl.VAL_F64), (64, fl.KERNEL_WIN, fl.VAL_BF16X2)]


@pytest.mark.parametrize("which", ["powerlaw", "stencil"])
def test_full_size_spmv_parity(gpu, request, which):
    m, n, cp, ri, v, x = request.getfixturevalue(which)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    mats = {W: fl.FastLoadMatrix(m, n, cp, ri, v, width=W) for W in (32, 64)}
    for W, kernel, kind in SCHEDULES:
        y = device_y(mats[W], x, kernel, kind, bt=256, R=8192)
        err = rel_err(y, y_or, scale)
        assert err <= TOL[kind], (which, W, kernel, kind, err)
    # FastLoad_CPU over the same 32-wide tiles agrees with the FP64 GPU path
    t = ob.pipeline(m, n, cp, ri, v, x, 32)
    y_cpu = ob.fastload_cpu_f64(t, t["sortx"], m)
    y_gpu = device_y(mats[32], x, fl.KERNEL_DROPIN, fl.VAL_F64)
    assert rel_err(y_gpu, y_cpu, scale) <= 1e-12


@pytest.mark.parametrize("which", ["powerlaw", "stencil"])
def test_full_size_learned_schedule(gpu, request, which):
    """The schedule the selector picks for the workload, run as the benchmark runs it."""
    m, n, cp, ri, v, x = request.getfixturevalue(which)
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    c, p, f = sel.Selector().choose(m, n, dcp, dri)
    run = sel.ScheduledSpMV(m, n, dcp, dri, dv, c)
    run.set_x(DeviceArray.from_host(x.astype(np.float32)))
    dy = DeviceArray(m, np.float32)
    run.step(dy)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x.astype(np.float32))
    assert rel_err(dy.to_host().astype(np.float64), y_or, scale) <= 1e-5, sel.SCHEDULES[c].name


def test_full_size_hyb_all_kinds(gpu, powerlaw):
    """The HYB schedule at full size, where its three kernels read their value and index streams
    with the streaming hint (about 490 MB per SpMV, above the 256 MB Infinity Cache): FP64 with
    19,456 heavy columns, FP32 and packed light values with 32,768, against the FP64 oracle
    (packed: the sub-matrix's stated rounding bound), y overwritten under assign."""
    m, n, cp, ri, v, x = powerlaw
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    y64, s64 = ob.spmv_oracle(m, n, cp, ri, v, x)
    M.prepare_hyb(19456, 128, 16384)
    assert rel_err(device_y_hyb(M, x, fl.VAL_F64), y64, s64) <= 1e-12
    M.prepare(fl.VAL_F32)
    M.prepare(fl.VAL_PACKED)
    M.prepare_hyb(None, 128, 16384)
    assert M.hyb_info()[0] == 32768
    y32, s32 = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x.astype(np.float32))
    assert rel_err(device_y_hyb(M, x, fl.VAL_F32), y32, s32) <= 1e-5
    yx, sx = ob.spmv_oracle(m, n, cp, ri, v, x.astype(np.float32).astype(np.float64))
    mb = M.pack_info(fl.KERNEL_HYB)[0]
    assert rel_err(device_y_hyb(M, x, fl.VAL_PACKED), yx, sx) <= 2.0 ** -(mb + 1) * (1 + 1e-6) + 1e-6


def device_y_hyb(M, x, kind):
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    dx = DeviceArray.from_host(np.asarray(x, dt))
    dsx = DeviceArray(M.n, dt)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.full(M.m, 9.0, dt))
    M.spmv(dsx, dy, fl.KERNEL_HYB, kind, assign=True)
    return dy.to_host().astype(np.float64)


def test_full_size_checksum_and_linearity(gpu, powerlaw):
    """1'(A x) = (1'A) x, with the column sums 1'A taken from the CSC input in FP64; and
    A(x1 + 2 x2) = A x1 + 2 A x2 in FP64, for the PB and drop-in schedules."""
    m, n, cp, ri, v, x = powerlaw
    colsum = np.add.reduceat(np.append(v, 0.0), cp[:-1]) * (np.diff(cp) > 0)
    x2 = fl.gen_vector(fl.VEC_UNIFORM, 44, n)
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    W32 = fl.FastLoadMatrix(m, n, cp, ri, v, width=32)
    for Mx, kernel in ((M, fl.KERNEL_PB), (W32, fl.KERNEL_DROPIN)):
        y1 = device_y(Mx, x, kernel, fl.VAL_F64, bt=256, R=8192)
        y2 = device_y(Mx, x2, kernel, fl.VAL_F64, bt=256, R=8192)
        y12 = device_y(Mx, x + 2 * x2, kernel, fl.VAL_F64, bt=256, R=8192)
        want = float(np.dot(colsum, x))
        bound = float(np.dot(np.abs(colsum), np.abs(x)))
        assert abs(y1.sum() - want) <= 1e-12 * bound * 10, kernel
        _, s1 = ob.spmv_oracle(m, n, cp, ri, v, np.abs(x) + 2 * np.abs(x2))
        assert rel_err(y12, y1 + 2 * y2, s1) <= 1e-12, kernel


@pytest.fixture(scope="module")
def uniform_42m():
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_UNIFORM, 7, 1 << 20, 1 << 20, (4e-5,))
    nnz = int(cp[n])
    assert nnz >= 1 << 25 and m * 8 <= nnz  # large enough for the two-chunk PB
    return m, n, cp, ri, v, fl.gen_vector(fl.VEC_NORMAL, 8, n)


@pytest.mark.parametrize("bt,R", [(64, 1024), (128, 4096), (256, 16384)])
def test_pb_two_chunks_parity(gpu, uniform_42m, bt, R):
    """PB on a matrix of >= 2^25 entries runs in two chunks of batches (chunk boundaries at
    multiples of 8 batches, per-chunk band tables, products at the front of the product array):
    y against the FP64 oracle for every batch size and band height, and fl_spmv_assign (first
    chunk writes y, second adds) equal bit for bit to y += A x on a cleared y."""
    m, n, cp, ri, v, x = uniform_42m
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    for kind in (fl.VAL_F32, fl.VAL_F64, fl.VAL_BF16X2):
        y = device_y(M, x, fl.KERNEL_PB, kind, bt=bt, R=R)
        assert rel_err(y, y_or, scale) <= TOL[kind], (bt, R, kind)
    dx = DeviceArray.from_host(x.astype(np.float32))
    dsx = DeviceArray(n, np.float32)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.full(m, 7.0, np.float32))
    M.spmv(dsx, dy, fl.KERNEL_PB, fl.VAL_F32, assign=True)
    ya = dy.to_host()
    dy0 = DeviceArray.zeros(m, np.float32)
    M.spmv(dsx, dy0, fl.KERNEL_PB, fl.VAL_F32)
    assert np.array_equal(ya, dy0.to_host())


def test_pb_two_chunks_graph_replays(gpu, uniform_42m):
    """fl_spmv_assign with the two-chunk PB (four kernels) captured in a hipGraph: every replay,
    with eager work between replays, equals the eager result to the last FP32 bit or so (the
    FP64 LDS sums of a band may add in another order)."""
    m, n, cp, ri, v, x = uniform_42m
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_F32)
    M.prepare_schedule(fl.KERNEL_PB, 128, 16384)
    s = fl.Stream()
    dx = DeviceArray.from_host(x.astype(np.float32))
    dsx = DeviceArray(n, np.float32)
    M.permute_x(dx, dsx)
    dy = DeviceArray(m, np.float32)

    def step():
        M.spmv(dsx, dy, fl.KERNEL_PB, fl.VAL_F32, stream=s, assign=True)
    step()
    s.synchronize()
    eager = dy.to_host()
    g = fl.Graph(s)
    with g:
        step()
    for rep in range(3):
        dy.copy_from(DeviceArray.from_host(np.full(m, 7.0, np.float32)))
        fl.synchronize()
        g.launch()
        s.synchronize()
        got = dy.to_host()
        assert np.allclose(got, eager, rtol=1e-6, atol=1e-6), rep
        step()
        s.synchronize()
