"""The batched configuration at its stated size on the GPU: 512 matrices (rows log-uniform in
1k-200k, column counts Zipf(2.0), ~150 M entries in all, seed 512), stacked block-diagonally.

  k = 1   the schedule the learned selector picks for the stacked matrix: the whole y against the
          compensated FP64 oracle (FP32 bar, 1e-5 relative to sum_j |a_rj x_j|)
  k = 8, 32   the row kernel: every block of the stacked result equals the same matrix run alone,
          bit for bit, and 4 sampled right-hand-side columns match the oracle
  graph   one replay of the captured k = 1 step equals the eager step
"""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray, Graph, Stream
from fastload_amd import workloads as wl

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


@pytest.fixture(scope="module")
def batch512(gpu):
    mats = wl.batch(512, 512)
    M_, N_, cp, ri, v, ro, co = fl.batch_assemble(mats)
    nnz = int(cp[-1])
    print(f"batch: {len(mats)} matrices, {M_} rows, {nnz} entries", flush=True)
    assert len(mats) == 512 and 1.2e8 < nnz < 1.8e8
    return mats, M_, N_, cp, ri, v, ro, co


def test_batch512_k1_learned_schedule(batch512):
    from fastload_amd import selector as sel
    mats, M_, N_, cp, ri, v, ro, co = batch512
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    c, prob, _ = sel.Selector().choose(M_, N_, dcp, dri)
    print(f"selector: {sel.SCHEDULES[c].name} p={prob:.3f}", flush=True)
    run = sel.ScheduledSpMV(M_, N_, dcp, dri, dv, c)
    del dcp, dri, dv
    x = np.random.default_rng(1).standard_normal(N_).astype(np.float32)
    run.set_x(DeviceArray.from_host(x))
    dy = DeviceArray(M_, np.float32)
    run.step(dy)
    y = dy.to_host().astype(np.float64)
    y_or, scale = ob.spmv_oracle(M_, N_, cp, ri, v.astype(np.float32), x)
    err = rel_err(y, y_or, scale)
    print(f"k=1 {sel.SCHEDULES[c].name}: rel err {err:.2e}", flush=True)
    assert err <= TOL32
    # one graph replay of the same step equals the eager result
    s = Stream()
    g = Graph(s)
    with g:
        run.step(dy, stream=s)
    y_eager = dy.to_host()
    dy.zero()
    g.launch()
    s.synchronize()
    assert np.array_equal(dy.to_host(), y_eager)


@pytest.mark.parametrize("k", [8, 32])
def test_batch512_row_kernel(batch512, k):
    mats, M_, N_, cp, ri, v, ro, co = batch512
    B = fl.FastLoadMatrix(M_, N_, cp, ri, v, width=64)
    B.prepare_rows()
    X = np.random.default_rng(k).standard_normal((N_, k)).astype(np.float32)
    dX = DeviceArray.from_host(X.ravel())
    dY = DeviceArray(M_ * k, np.float32)
    B.spmm(dX, dY, k, fl.VAL_F32)
    Y = dY.to_host().reshape(M_, k)
    del B, dX, dY
    # every block equals its matrix run alone, bit for bit
    for i, (m, n, c, r, vv) in enumerate(mats):
        Q = fl.FastLoadMatrix(m, n, c, r, vv, width=64)
        Q.prepare_rows()
        dXi = DeviceArray.from_host(np.ascontiguousarray(X[co[i]:co[i + 1]]).ravel())
        dYi = DeviceArray(m * k, np.float32)
        Q.spmm(dXi, dYi, k, fl.VAL_F32)
        Yi = dYi.to_host().reshape(m, k)
        assert np.array_equal(Yi.view(np.uint32), Y[ro[i]:ro[i + 1]].view(np.uint32)), i
    print(f"k={k}: 512 blocks bit-identical to per-matrix runs", flush=True)
    # sampled right-hand sides against the oracle
    for j in np.random.default_rng(100 + k).choice(k, 4, replace=False):
        y_or, scale = ob.spmv_oracle(M_, N_, cp, ri, v.astype(np.float32), np.ascontiguousarray(X[:, j]))
        err = rel_err(Y[:, j].astype(np.float64), y_or, scale)
        print(f"k={k} column {j}: rel err {err:.2e}", flush=True)
        assert err <= TOL32
