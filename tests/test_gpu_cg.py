"""Conjugate gradients on the resident matrix (fl_cg_*) against a numpy FP64 CG on the same
symmetric positive definite systems: a 3-D 7-point Laplacian and a diagonally dominant
structurally symmetric matrix, through several schedules, in FP64 and FP32, with the iterations
replayed from a hipGraph (a second solve reuses the captured graph)."""
import numpy as np
import pytest
import scipy.sparse as sp

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu


def laplacian3d(g):
    one = sp.diags([-np.ones(g - 1), 2 * np.ones(g), -np.ones(g - 1)], [-1, 0, 1])
    eye = sp.identity(g)
    A = (sp.kron(sp.kron(one, eye), eye) + sp.kron(sp.kron(eye, one), eye) + sp.kron(sp.kron(eye, eye), one))
    return sp.csc_matrix(A)


def dominant_symgeom():
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_SYMGEOM, 1138, 1138, 1138)
    A = sp.csc_matrix((v, ri, cp), shape=(m, n))
    A = A - sp.diags(A.diagonal())
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    return sp.csc_matrix(A)


def numpy_cg(A, b, rtol, max_iter):
    x = np.zeros_like(b)
    r = b - A @ x
    p = r.copy()
    rr = r @ r
    bn = np.linalg.norm(b)
    for k in range(max_iter):
        q = A @ p
        alpha = rr / (p @ q)
        x += alpha * p
        r -= alpha * q
        rn = r @ r
        if np.sqrt(rn) / bn <= rtol:
            return x, k + 1
        p = r + (rn / rr) * p
        rr = rn
    return x, max_iter


SCHEDULES = [("dropin_w32", 32, fl.KERNEL_DROPIN, 0, 0), ("pb", 64, fl.KERNEL_PB, 128, 4096),
             ("winr", 64, fl.KERNEL_WINR, 64, 0), ("row", 64, fl.KERNEL_ROW, 0, 0), ("fast", 64, fl.KERNEL_FAST, 0, 0)]


@pytest.mark.parametrize("kind", [fl.VAL_F64, fl.VAL_F32], ids=["f64", "f32"])
@pytest.mark.parametrize("sched", SCHEDULES, ids=[s[0] for s in SCHEDULES])
@pytest.mark.parametrize("system", ["laplacian24", "symgeom"])
def test_cg_matches_numpy(gpu, system, sched, kind):
    A = laplacian3d(24) if system == "laplacian24" else dominant_symgeom()
    A.sort_indices()
    n = A.shape[0]
    name, W, kernel, bt, R = sched
    M = fl.FastLoadMatrix(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, width=W)
    if kernel == fl.KERNEL_ROW:
        M.prepare_rows()
    else:
        if kind == fl.VAL_F32:
            M.prepare(fl.VAL_F32)
        if kernel in (fl.KERNEL_PB, fl.KERNEL_WINR):
            M.prepare_schedule(kernel, bt, R)
    dt = np.float64 if kind == fl.VAL_F64 else np.float32
    rtol = 1e-10 if kind == fl.VAL_F64 else 1e-5
    cg = fl.CG(M, kernel, kind)
    rng = np.random.default_rng(7)
    for trial in range(2):  # the second solve replays the graph captured by the first
        b = rng.standard_normal(n)
        x_np, it_np = numpy_cg(A, b, rtol, 2000)
        db = DeviceArray.from_host(b.astype(dt))
        dx = DeviceArray.zeros(n, dt)
        it, res = cg.solve(db, dx, max_iter=2000, rtol=rtol, check_every=8)
        x = dx.to_host().astype(np.float64)
        true_res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print(f"{system} {name} kind={kind} trial={trial}: {it} iterations (numpy {it_np}), "
              f"res {res:.2e}, true res {true_res:.2e}", flush=True)
        assert res <= rtol and it < 2000
        if kind == fl.VAL_F64:
            assert abs(it - it_np) <= 8 + 2
            assert true_res <= 10 * rtol
            assert np.linalg.norm(x - x_np) / np.linalg.norm(x_np) <= 1e-7
        else:
            assert true_res <= 1e-3
            assert np.linalg.norm(x - x_np) / np.linalg.norm(x_np) <= 1e-2


def test_cg_zero_rhs_and_refusals(gpu):
    A = laplacian3d(6)
    n = A.shape[0]
    M = fl.FastLoadMatrix(n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, width=64)
    cg = fl.CG(M, fl.KERNEL_FAST, fl.VAL_F64)
    db = DeviceArray.zeros(n, np.float64)
    dx = DeviceArray.from_host(np.ones(n))
    assert cg.solve(db, dx)[0] == 0
    with pytest.raises(fl.FastLoadError):  # PB not prepared: refused at creation
        fl.CG(M, fl.KERNEL_PB, fl.VAL_F64)
    R = fl.FastLoadMatrix(n, n + 1, np.append(A.indptr, A.indptr[-1]).astype(np.int32), A.indices.astype(np.int32),
                          A.data, width=64)
    with pytest.raises(fl.FastLoadError):  # not square
        fl.CG(R, fl.KERNEL_FAST, fl.VAL_F64)
