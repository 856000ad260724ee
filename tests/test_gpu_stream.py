"""The streaming cache hint changes how the WIN/WINR, PB and HYB schedules load their read-once
streams, never what they compute: y under FL_STREAM_ALWAYS equals y under FL_STREAM_NEVER up to
the order of the schedules' FP64 atomic adds (which is not fixed from one call to the next: a
rounding of the FP64 sum, or one unit in the last place of an FP32 y), for every schedule and
value kind, and both agree with the FP64 oracle."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

KINDS = {fl.KERNEL_WIN: (fl.VAL_F64, fl.VAL_F32, fl.VAL_BF16, fl.VAL_BF16X2, fl.VAL_PACKED),
         fl.KERNEL_WINR: (fl.VAL_F32, fl.VAL_PACKED),
         fl.KERNEL_PB: (fl.VAL_F64, fl.VAL_F32, fl.VAL_BF16, fl.VAL_BF16X2, fl.VAL_PACKED),
         fl.KERNEL_HYB: (fl.VAL_F64, fl.VAL_F32, fl.VAL_PACKED)}


def _run(M, x, kernel, kind):
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    dx = DeviceArray.from_host(np.asarray(x, dt))
    dsx = DeviceArray(M.n, dt)
    M.permute_x(dx, dsx)
    yt = np.float64 if kind in (fl.VAL_F64, fl.VAL_BF16X2) else np.float32
    dy = DeviceArray.from_host(np.full(M.m, 3.0, yt))
    M.spmv(dsx, dy, kernel, kind, assign=True)
    return dy.to_host()


@pytest.mark.parametrize("kernel", [fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB, fl.KERNEL_HYB])
def test_stream_policy_does_not_change_y(gpu, kernel):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 61, 90000, 70000, (1.8, 6000))
    x = np.random.default_rng(61).standard_normal(n)
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    for kind in KINDS[kernel]:
        if kind != fl.VAL_F64:
            M.prepare(kind)
    if kernel == fl.KERNEL_HYB:
        M.prepare_hyb(4096, 128, 16384)
    else:
        M.prepare_schedule(kernel, 128, 8192)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    try:
        for kind in KINDS[kernel]:
            fl.set_stream_policy(fl.STREAM_NEVER)
            y0 = _run(M, x, kernel, kind)
            fl.set_stream_policy(fl.STREAM_ALWAYS)
            y1 = _run(M, x, kernel, kind)
            same = 1e-12 if y0.dtype == np.float64 else 3e-7
            assert rel_err(y1.astype(np.float64), y0.astype(np.float64), scale) <= same, (kernel, kind)
            tol = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5, fl.VAL_BF16X2: 1e-4}.get(kind, 1e-2)
            assert rel_err(y1.astype(np.float64), y_or, scale) <= tol, (kernel, kind)
    finally:
        fl.set_stream_policy(fl.STREAM_AUTO)


def test_preparation_on_two_user_streams(gpu):
    """Layout, PB and HYB preparation of two matrices enqueued on two streams of the caller, one
    after the other without waiting: their temporaries come from the stream-ordered pool, so each
    must only reuse memory its own stream has finished with. Both y agree with the oracle."""
    streams = [fl.Stream(), fl.Stream()]
    specs = [(71, 60000, 50000), (72, 80000, 65000)]
    mats, data = [], []
    for (seed, m0, n0), s in zip(specs, streams):
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, seed, m0, n0, (1.8, 5000))
        M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64, stream=s)
        M.prepare(fl.VAL_F32, stream=s)
        M.prepare_schedule(fl.KERNEL_PB, 128, 8192, stream=s)
        mats.append(M)
        data.append((m, n, cp, ri, v))
    for M, s in zip(mats, streams):
        M.prepare_hyb(4096, 128, 16384, stream=s)
    for (m, n, cp, ri, v), M, s in zip(data, mats, streams):
        x = np.random.default_rng(m).standard_normal(n).astype(np.float32)
        dx = DeviceArray.from_host(x)
        dsx = DeviceArray(n, np.float32)
        M.permute_x(dx, dsx, stream=s)
        y_or, scale = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x)
        for kernel in (fl.KERNEL_PB, fl.KERNEL_HYB):
            dy = DeviceArray(m, np.float32)
            M.spmv(dsx, dy, kernel, fl.VAL_F32, stream=s, assign=True)
            s.synchronize()
            assert rel_err(dy.to_host().astype(np.float64), y_or, scale) <= 1e-5, kernel
