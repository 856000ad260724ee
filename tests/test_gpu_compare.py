"""The measurement-only comparators (libfastload_compare.so) on the GPU: the reference's naive CSC
kernel (src/CSCSpMV.h:4-27) and rocSPARSE's CSR SpMV (in place of src/cusparse_cuda.h), against
the FP64 oracle, and their reference-style host entry points on 1138_bus with the integer data of
src/main.cu:68-69,93-96 (exact)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_bind as ob
from conftest import GOLDEN
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-5}


def _mats():
    yield "zipf", fl.gen_matrix(fl.GEN_ZIPF, 9, 30000, 50000, (1.8, 2000))
    yield "stencil", fl.gen_matrix(fl.GEN_STENCIL27, 9, 0, 0, (16,))
    yield "empty_cols", (4, 6, np.array([0, 0, 2, 2, 3, 3, 3], np.int32), np.array([1, 3, 0], np.int32),
                         np.array([1., 2, 3]))


@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("name,mat", list(_mats()), ids=[c[0] for c in _mats()])
def test_comparators_match_oracle(gpu, name, mat, dt):
    from fastload_amd import compare as cmpl
    m, n, cp, ri, v = mat
    x = np.random.default_rng(5).standard_normal(n).astype(dt)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v.astype(dt), x)
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v.astype(dt))
    dx = DeviceArray.from_host(x)
    dy = DeviceArray.zeros(max(m, 1), dt)
    cmpl.csc_naive_device(m, n, dcp, dri, dv, dx, dy)
    assert rel_err(dy.to_host()[:m].astype(np.float64), y_or, scale) <= TOL[dt]
    A = sp.csc_matrix((v.astype(dt), ri, cp), shape=(m, n)).tocsr()
    A.sort_indices()
    R = cmpl.RocsparseCSR(m, n, DeviceArray.from_host(A.indptr.astype(np.int32)),
                          DeviceArray.from_host(A.indices.astype(np.int32)), DeviceArray.from_host(A.data), None)
    dy2 = DeviceArray.from_host(np.full(max(m, 1), 9.0, dt))
    R.run(dx, dy2)
    fl.synchronize()
    assert rel_err(dy2.to_host()[:m].astype(np.float64), y_or, scale) <= TOL[dt]


def test_reference_style_comparators_1138(gpu, capfd):
    from fastload_amd import compare as cmpl
    m, n, nnz, _, rp, ci, v = fl.mmio_allinone(str(GOLDEN / "1138_bus.mtx"))
    v = (np.arange(nnz) % 10).astype(np.float64)
    x = (np.arange(n) % 10).astype(np.float64)
    A = sp.csr_matrix((v, ci, rp), shape=(m, n))
    y_gold = A @ x
    C = A.tocsc()
    C.sort_indices()
    y, ms, ok = cmpl.csc_spmv(m, n, C.indptr, C.indices, C.data, x, y_golden=y_gold, filename="1138_bus")
    assert ok and np.array_equal(y, y_gold) and ms > 0
    y2, ms2, ok2 = cmpl.rocsparse_spmv(m, n, rp, ci, v, x, y_result=y_gold)
    assert ok2 and np.array_equal(y2, y_gold)
    out = capfd.readouterr().out
    assert "BASIC CSC GPU PASS!" in out and "rocSPARSE test PASSED!" in out
