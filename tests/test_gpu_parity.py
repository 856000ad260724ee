"""GPU parity: the HIP path (through the C ABI) against the oracle and the reference golden
fixtures. Preprocessing must be bit-exact; y must match the FP64 oracle within 1e-12 relative
(FP64 mode), 1e-5 (FP32 mode), 1e-4 (split hi+lo BF16) and 2^-7 (BF16 alone, error reported),
relative to sum_j |a_rj x_j| per row; integer data must be exact in FP64."""
import json

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_bind as ob
from conftest import GOLDEN
from gpu_util import device_y, export_matches_oracle, rel_err

import fastload_amd as fl

pytestmark = pytest.mark.gpu

TOL = {fl.VAL_F64: 1e-12, fl.VAL_F32: 1e-5, fl.VAL_BF16X2: 1e-4, fl.VAL_BF16: 2.0 ** -7}


def golden_1138():
    return np.load(GOLDEN / "ref_layout_1138_bus.npz")


# ------------------------------------------------------------ reference host API path
def test_host_api_pipeline_1138_bus(gpu):
    """mmio_allinone -> csr_to_csc -> ColSort -> formatTransform -> FastLoad_spmv, as src/main.cu."""
    g = golden_1138()
    m, n, nnz, sym, rp, ci, v = fl.mmio_allinone(str(GOLDEN / "1138_bus.mtx"))
    assert (m, n, nnz) == (1138, 1138, 4054)
    v = (np.arange(nnz) % 10).astype(np.float64)  # src/main.cu:68-69
    cv, cr, cp, npc, _ = fl.csr_to_csc(m, n, v, rp, ci)
    assert np.array_equal(cp, g["csc_colptr"]) and np.array_equal(cr, g["csc_rowidx"]) and np.array_equal(cv, g["csc_val"])
    assert np.array_equal(npc, np.diff(g["csc_colptr"]))
    x = (np.arange(n) % 10).astype(np.float64)  # src/main.cu:93-96
    tr, tv, sn, sx, _ = fl.ColSort(m, n, npc, cp, cr, cv, x)
    assert np.array_equal(sn, g["sortnnz"]) and np.array_equal(sx, g["sortx"])
    assert np.array_equal(tr, g["tmp_row"]) and np.array_equal(tv, g["tmp_val"])
    tiles, h_count, _, _ = fl.formatTransform(tr, tv, sn, nnz, m, n)
    assert h_count == 14 and tiles.tilenum == 139
    for k in ["tile_ptr", "tile_len", "tile_colidx", "tile_format", "sortrowidx", "sortval", "sortrowindex",
              "countrow", "segmentoffset"]:
        assert np.array_equal(getattr(tiles, k), g[f"w32_{k}"]), k
    y_golden = ob.spmv_csr(m, rp, ci, v, x)
    y, ms, passed = fl.FastLoad_spmv(tiles, nnz, n, m, sx, y_golden, "1138_bus")
    assert passed and np.array_equal(y, y_golden) and ms > 0


def test_host_api_csr2csc_random(gpu):
    rng = np.random.default_rng(11)
    for (m, n, d) in [(1, 1, 1.0), (37, 1000, 0.01), (2000, 3, 0.5), (5000, 4000, 0.002)]:
        A = sp.random(m, n, density=d, random_state=rng, format="csr")
        A.sort_indices()
        cv, cr, cp, npc, _ = fl.csr_to_csc(m, n, A.data, A.indptr, A.indices)
        rp2, ci2 = A.indptr.astype(np.int32), A.indices.astype(np.int32)
        e = ob.csr2csc(m, n, rp2, ci2, A.data)
        assert np.array_equal(cp, e[0]) and np.array_equal(cr, e[1]) and np.array_equal(cv, e[2])


def test_host_api_invalid_inputs(gpu):
    cp = np.array([0, 1, 2], np.int32)
    with pytest.raises(fl.FastLoadError):  # nnzPerCol disagrees with colptr
        fl.ColSort(2, 2, np.array([2, 0], np.int32), cp, np.array([0, 1], np.int32), np.ones(2), np.ones(2))
    with pytest.raises(fl.FastLoadError):  # unsorted lengths
        fl.formatTransform(np.zeros(3, np.int32), np.zeros(3), np.array([2, 1], np.int32), 3, 2, 2)
    with pytest.raises(fl.FastLoadError):  # row index out of range
        fl.FastLoadMatrix(2, 2, cp, np.array([0, 5], np.int32), np.ones(2))


# ------------------------------------------------------------------ device-resident path
def _cases():
    rng = np.random.default_rng(5)
    cases = []
    g = golden_1138()
    cases.append(("1138_bus", int(g["m"]), int(g["n"]), g["csc_colptr"], g["csc_rowidx"], g["csc_val"], g["x"]))
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_SYMGEOM, 1138, 1138, 1138)
    cases.append(("symgeom", m, n, cp, ri, v, fl.gen_vector(fl.VEC_UNIFORM, 1139, n)))
    for name, fam, mm, nn, pr in [("zipf", fl.GEN_ZIPF, 3000, 5000, (1.8, 800)), ("banded", fl.GEN_BANDED, 700, 700, (37,)),
                                  ("blockdiag", fl.GEN_BLOCKDIAG, 4096, 4096, (64, 0)),
                                  ("uniform", fl.GEN_UNIFORM, 900, 20000, (3e-3,)),
                                  ("zipfdense", fl.GEN_ZIPF_DENSE, 5000, 3000, (2.0, 100, 0.01, 2000))]:
        m, n, cp, ri, v = fl.gen_matrix(fam, 7, mm, nn, pr)
        cases.append((name, m, n, cp, ri, v, rng.standard_normal(n)))
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_STENCIL27, 9, 0, 0, (12,))
    cases.append(("stencil12", m, n, cp, ri, v, rng.standard_normal(n)))
    # edge cases: one column; empty columns; a column per row count; rectangular; all-empty
    cases.append(("n1", 5, 1, np.array([0, 3], np.int32), np.array([0, 2, 4], np.int32), np.array([1., 2, 3]), np.ones(1)))
    cases.append(("empty_cols", 4, 6, np.array([0, 0, 2, 2, 3, 3, 3], np.int32), np.array([1, 3, 0], np.int32),
                  np.array([1., 2, 3]), np.arange(6.0)))
    cases.append(("all_empty", 3, 4, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0), np.ones(4)))
    dense = np.arange(1.0, 1 + 70 * 130).reshape(70, 130) % 7  # dense-ish: full 32/64-tiles of one row
    A = sp.csc_matrix(dense)
    A.sort_indices()
    cases.append(("dense70x130", 70, 130, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, rng.standard_normal(130)))
    return cases


CASES = _cases()


@pytest.mark.parametrize("W", [32, 64])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_layout_bit_exact(gpu, case, W):
    name, m, n, cp, ri, v, x = case
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    bad, t = export_matches_oracle(M, m, n, cp, ri, v, x, W)
    assert not bad, f"{name} W={W}: arrays differ: {bad}"


def test_device_layout_matches_reference_golden(gpu):
    for fname in ("ref_layout_1138_bus.npz", "ref_layout_symgeom.npz"):
        g = np.load(GOLDEN / fname)
        for W in (32, 64):
            M = fl.FastLoadMatrix(int(g["m"]), int(g["n"]), g["csc_colptr"], g["csc_rowidx"], g["csc_val"], width=W)
            e = M.export()
            keys = ["tile_ptr", "tile_len", "tile_colidx", "sortrowidx", "sortval", "sortrowindex", "countrow",
                    "segmentoffset"] + (["tile_format"] if W == 32 else [])
            for k in keys:
                assert np.array_equal(e[k], g[f"w{W}_{k}"]), (fname, W, k)
            assert np.array_equal(e["perm"], g["perm"])


KERNELS = [(fl.KERNEL_DROPIN, fl.VAL_F64, 32), (fl.KERNEL_DROPIN, fl.VAL_F64, 64), (fl.KERNEL_DROPIN, fl.VAL_F32, 32),
           (fl.KERNEL_DROPIN, fl.VAL_F32, 64), (fl.KERNEL_FAST, fl.VAL_F64, 64), (fl.KERNEL_FAST, fl.VAL_F32, 64),
           (fl.KERNEL_FAST, fl.VAL_BF16, 64), (fl.KERNEL_FAST, fl.VAL_BF16X2, 64)]
KERNELS += [(k, v, 64) for k in (fl.KERNEL_WIN, fl.KERNEL_PB, fl.KERNEL_WINR)
            for v in (fl.VAL_F64, fl.VAL_F32, fl.VAL_BF16, fl.VAL_BF16X2)]
KNAMES = ["dropin", "fast", "win", "pb", "winr"]
KIDS = [f"{KNAMES[k]}-{['f64', 'f32', 'bf16', 'bf16x2'][v]}-w{w}" for k, v, w in KERNELS]


@pytest.mark.parametrize("kernel,kind,W", KERNELS, ids=KIDS)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_spmv_parity(gpu, case, kernel, kind, W):
    name, m, n, cp, ri, v, x = case
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    y = device_y(M, x, kernel, kind)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    err = rel_err(y, y_or, scale)
    assert err <= TOL[kind], f"{name}: rel err {err:.3e} > {TOL[kind]:.1e}"
    if kind == fl.VAL_F64:
        # FastLoad_CPU (same tiles, serial order) agrees to the same tolerance
        t = ob.pipeline(m, n, cp, ri, v, x, W)
        y_cpu = ob.fastload_cpu_f64(t, t["sortx"], m)
        assert rel_err(y, y_cpu, scale) <= 1e-12


@pytest.mark.parametrize("kernel,kind,W", [k for k in KERNELS if k[1] in (fl.VAL_F64, fl.VAL_F32)],
                         ids=[i for i, k in zip(KIDS, KERNELS) if k[1] in (fl.VAL_F64, fl.VAL_F32)])
def test_spmv_integer_data_exact(gpu, kernel, kind, W):
    """The reference's exactness trick: values i%10, x i%10 -> every partial sum is an exact small
    integer, so any summation order gives the same bits (src/main.cu:68-69,93-96, FastLoad_CPU.h:33-45)."""
    g = golden_1138()
    m, n = int(g["m"]), int(g["n"])
    M = fl.FastLoadMatrix(m, n, g["csc_colptr"], g["csc_rowidx"], g["csc_val"], width=W)
    y = device_y(M, g["x"], kernel, kind)
    assert np.array_equal(y, g["w32_y"])


def test_kat_9x9_gpu(gpu):
    kat = json.loads((GOLDEN / "kat_9x9.json").read_text())
    A = sp.csc_matrix(np.array(kat["A"], np.float64))
    A.sort_indices()
    for W in (32, 64):
        M = fl.FastLoadMatrix(9, 9, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, width=W)
        for kernel, kind, w in KERNELS:
            if w != W or kind in (fl.VAL_BF16,):
                continue
            y = device_y(M, np.ones(9), kernel, kind)
            assert y.tolist() == kat["row_sums"], (W, kernel, kind)


@pytest.mark.parametrize("kernel", [fl.KERNEL_WIN, fl.KERNEL_PB, fl.KERNEL_WINR], ids=["win", "pb", "winr"])
@pytest.mark.parametrize("bt,R", [(64, 1024), (128, 2048), (256, 4096), (256, 16384)])
def test_batched_schedule_params(gpu, kernel, bt, R):
    """Multi-batch / multi-band coverage on matrices small enough to check against the oracle."""
    for name, m, n, cp, ri, v, x in CASES:
        if name in ("all_empty",):
            continue
        M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
        y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
        for kind in (fl.VAL_F64, fl.VAL_F32):
            y = device_y(M, x, kernel, kind, bt=bt, R=R)
            err = rel_err(y, y_or, scale)
            assert err <= TOL[kind], f"{name} bt={bt} R={R} kind={kind}: {err:.3e}"


def test_reference_style_driver(gpu, tmp_path):
    """The drop-in boundary from C++: a src/main.cu-style driver compiled against
    include/fastload_compat.hpp (the reference's own function names and signatures)."""
    import subprocess
    from conftest import ROOT
    exe = tmp_path / "drop_in_main"
    subprocess.run(["g++", "-std=c++17", "-I", str(ROOT / "include"), str(ROOT / "tests" / "drop_in_main.cpp"),
                    "-L", str(ROOT / "fastload_amd"), "-lfastload", "-lfastload_compare",
                    f"-Wl,-rpath,{ROOT / 'fastload_amd'}", "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe), str(GOLDEN / "1138_bus.mtx")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Check FASTLOAD GPU PASS!" in r.stdout, r.stdout + r.stderr
    assert "Check NO PASS" not in r.stdout, r.stdout  # FastLoad_cpu and FastLoad_spmv both exact
    assert "BASIC CSC GPU PASS!" in r.stdout and "rocSPARSE test PASSED!" in r.stdout, r.stdout


@pytest.mark.parametrize("kernel,kind,W", [(fl.KERNEL_DROPIN, fl.VAL_F64, 32), (fl.KERNEL_FAST, fl.VAL_F32, 64),
                                           (fl.KERNEL_WIN, fl.VAL_F32, 64), (fl.KERNEL_PB, fl.VAL_F32, 64),
                                           (fl.KERNEL_PB, fl.VAL_BF16X2, 64), (fl.KERNEL_PB, fl.VAL_F64, 64),
                                           (fl.KERNEL_WINR, fl.VAL_F32, 64)],
                         ids=["dropin-f64", "fast-f32", "win-f32", "pb-f32", "pb-bf16x2", "pb-f64", "winr-f32"])
@pytest.mark.parametrize("name", ["zipf", "empty_cols", "all_empty", "uniform"])
def test_spmv_assign_overwrites_y(gpu, kernel, kind, W, name):
    """fl_spmv_assign computes y = A x whatever y held before (rows without entries become 0)."""
    from fastload_amd import DeviceArray
    _, m, n, cp, ri, v, x = next(c for c in CASES if c[0] == name)
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    if kernel in (fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB):
        M.prepare_schedule(kernel, 128, 1024)
    dbl = kind in (fl.VAL_F64, fl.VAL_BF16X2)
    if kind != fl.VAL_F64:
        M.prepare(kind)
    xdt = np.float64 if kind == fl.VAL_F64 else np.float32
    dx = DeviceArray.from_host(np.asarray(x, xdt))
    dsx = DeviceArray(max(n, 1), xdt)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.full(max(m, 1), 123.0, np.float64 if dbl else np.float32))
    M.spmv(dsx, dy, kernel, kind, assign=True)
    y = dy.to_host().astype(np.float64)[:m]
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    assert rel_err(y, y_or[:m], scale[:m]) <= TOL[kind]


def test_switching_batch_order_invalidates_the_other_schedule(gpu):
    """WIN and PB share the column-ordered batches; WINR regroups the tiles by row window, so
    preparing it retires the others until they are prepared again (and vice versa)."""
    name, m, n, cp, ri, v, x = next(c for c in CASES if c[0] == "banded")
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    assert rel_err(device_y(M, x, fl.KERNEL_WIN, fl.VAL_F64, bt=64), y_or, scale) <= TOL[fl.VAL_F64]
    assert rel_err(device_y(M, x, fl.KERNEL_WINR, fl.VAL_F64, bt=64), y_or, scale) <= TOL[fl.VAL_F64]
    from fastload_amd import DeviceArray
    dx = DeviceArray.from_host(np.asarray(x, np.float64))
    dsx = DeviceArray(n, np.float64)
    M.permute_x(dx, dsx)
    dy = DeviceArray.zeros(m, np.float64)
    with pytest.raises(fl.FastLoadError):
        M.spmv(dsx, dy, fl.KERNEL_WIN, fl.VAL_F64)  # WIN's batches were replaced by WINR's
    assert rel_err(device_y(M, x, fl.KERNEL_PB, fl.VAL_F64, bt=64, R=1024), y_or, scale) <= TOL[fl.VAL_F64]
    with pytest.raises(fl.FastLoadError):
        M.spmv(dsx, dy, fl.KERNEL_WINR, fl.VAL_F64)


def _random_case(seed):
    """A seeded random matrix: family, shape (m != n allowed, down to 1) and parameters drawn."""
    rng = np.random.default_rng(1000 + seed)
    m, n = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))
    fam = seed % 4
    if fam == 0:
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, seed, m, n, (float(rng.uniform(1.3, 2.5)), int(rng.integers(1, 500))))
    elif fam == 1:
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_UNIFORM, seed, m, n, (float(10 ** rng.uniform(-4, -1)),))
    elif fam == 2:
        k = max(1, min(m, n) // int(rng.integers(1, 8)))
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_BANDED, seed, k, k, (int(rng.integers(1, 40)),))
    else:
        m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF_DENSE, seed, m, n, (2.0, 50, 0.02, int(rng.integers(1, m + 1))))
    return m, n, cp, ri, v, rng.standard_normal(n)


@pytest.mark.parametrize("seed", range(12))
def test_random_shapes_layout_and_schedules(gpu, seed):
    """Seeded random shapes and families: the layout is bit-exact at both widths and every
    schedule family agrees with the FP64 oracle, at a random batch size and band height."""
    m, n, cp, ri, v, x = _random_case(seed)
    rng = np.random.default_rng(seed)
    y_or, scale = ob.spmv_oracle(m, n, cp, ri, v, x)
    for W in (32, 64):
        M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
        bad, _ = export_matches_oracle(M, m, n, cp, ri, v, x, W)
        assert not bad, (seed, W, bad)
        runs = [(fl.KERNEL_DROPIN, fl.VAL_F64), (fl.KERNEL_DROPIN, fl.VAL_F32)]
        if W == 64:
            runs += [(fl.KERNEL_FAST, fl.VAL_F32), (fl.KERNEL_WIN, fl.VAL_F32), (fl.KERNEL_WINR, fl.VAL_F32),
                     (fl.KERNEL_PB, fl.VAL_F32), (fl.KERNEL_PB, fl.VAL_F64), (fl.KERNEL_WIN, fl.VAL_BF16X2)]
        for kernel, kind in runs:
            bt = int(rng.choice([64, 128, 256]))
            R = int(rng.choice([1024, 4096, 16384]))
            y = device_y(M, x, kernel, kind, bt=bt, R=R)
            assert rel_err(y, y_or, scale) <= TOL[kind], (seed, m, n, W, kernel, kind, bt, R)


@pytest.mark.parametrize("case", ["win_unprepared", "pb_unprepared", "winr_unprepared", "dropin_bf16", "fast_w32",
                                  "f32_values_missing", "unknown_kernel"])
def test_refused_assign_leaves_y_unchanged(gpu, case):
    """A call the library refuses returns FL_ERR_INVALID before anything is enqueued: y keeps
    what it held (fl_spmv_assign clears y only once the call is known to be valid)."""
    from fastload_amd import DeviceArray
    _, m, n, cp, ri, v, x = next(c for c in CASES if c[0] == "zipf")
    W = 32 if case == "fast_w32" else 64
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=W)
    if case != "f32_values_missing":
        M.prepare(fl.VAL_F32)
    kernel, kind = {"win_unprepared": (fl.KERNEL_WIN, fl.VAL_F32), "pb_unprepared": (fl.KERNEL_PB, fl.VAL_F32),
                    "winr_unprepared": (fl.KERNEL_WINR, fl.VAL_F32), "dropin_bf16": (fl.KERNEL_DROPIN, fl.VAL_BF16),
                    "fast_w32": (fl.KERNEL_FAST, fl.VAL_F32), "f32_values_missing": (fl.KERNEL_DROPIN, fl.VAL_F32),
                    "unknown_kernel": (9, fl.VAL_F32)}[case]
    dsx = DeviceArray.from_host(np.asarray(x, np.float32))
    dy = DeviceArray.from_host(np.full(m, 123.0, np.float32))
    with pytest.raises(fl.FastLoadError) as e:
        M.spmv(dsx, dy, kernel, kind, assign=True)
    assert e.value.rc == -8
    assert np.all(dy.to_host() == 123.0)
