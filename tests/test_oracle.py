"""Pins the CPU oracle (oracle/fastload_oracle.c) before anything is checked against it.

Each check is chosen so a plausible mistake (dropped term, wrong index, unstable tie order,
transposed operand) fails: golden vectors from the reference's own CPU twins, independent
library routines (scipy, numpy, math.fsum), closed forms and brute force on tiny inputs.
"""
import json
import math

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle_bind as ob
from conftest import GOLDEN

MTX = str(GOLDEN / "1138_bus.mtx")


def csc_of(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


# ---------------------------------------------------------------------------- loader
def test_loader_1138_bus_facts():
    rc, m, n, nnz, sym, rp, ci, v = ob.mmio_allinone(MTX)
    assert rc == 0 and (m, n, nnz, sym) == (1138, 1138, 4054, 1)
    # 2,596 stored entries -> 4,054 after mirroring: 1,138 diagonal + 2 x 1,458 off-diagonal
    rows = np.repeat(np.arange(m), np.diff(rp))
    assert np.sum(rows == ci) == 1138


def test_loader_matches_scipy_mmread():
    rc, m, n, nnz, sym, rp, ci, v = ob.mmio_allinone(MTX)
    ref = scipy.io.mmread(MTX).tocsr()
    mine = sp.csr_matrix((v, ci, rp), shape=(m, n))
    assert (abs(mine - ref)).max() == 0


def test_loader_order_rule(tmp_path):
    # symmetric file: mirrors are inserted into row c when line (r, c) is read (AUX:180-203)
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n2 1 1.5\n1 1 2\n3 1 3\n3 2 4\n")
    rc, m, n, nnz, sym, rp, ci, v = ob.mmio_allinone(str(p))
    assert rc == 0 and nnz == 7
    assert rp.tolist() == [0, 3, 5, 7]
    assert ci.tolist() == [1, 0, 2, 0, 2, 0, 1]
    assert v.tolist() == [1.5, 2, 3, 1.5, 4, 3, 4]


def test_loader_matches_reference_twin():
    if not ob.ref_available(32):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    a = ob.mmio_allinone(MTX)
    b = ob.ref_mmio_allinone(MTX)
    assert a[:5] == b[:5]
    for x, y in zip(a[5:], b[5:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("text,code", [
    ("not a banner\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n", -3),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -5),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -6),
])
def test_loader_errors(tmp_path, text, code):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    assert ob.mmio_allinone(str(p))[0] == code


def test_loader_missing_file():
    assert ob.mmio_allinone("/nonexistent/x.mtx")[0] == -1


# ---------------------------------------------------------------------------- csr2csc
def test_csr2csc_matches_scipy():
    rng = np.random.default_rng(1)
    A = sp.random(300, 200, density=0.03, random_state=rng, format="csr")
    A.sort_indices()
    cp, ri, cv, npc = ob.csr2csc(300, 200, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    r = csc_of(A)
    assert np.array_equal(cp, r[0]) and np.array_equal(ri, r[1]) and np.array_equal(cv, r[2])
    assert np.array_equal(npc, np.diff(r[0])) and cp[-1] == A.nnz


def test_csr2csc_unsorted_rows_stable_by_row():
    # columns inside a CSR row may be unsorted (loader order); CSC rows must still ascend
    rp = np.array([0, 3, 5], np.int32)
    ci = np.array([2, 0, 1, 1, 0], np.int32)
    v = np.arange(5, dtype=np.float64)
    cp, ri, cv, npc = ob.csr2csc(2, 3, rp, ci, v)
    assert cp.tolist() == [0, 2, 4, 5] and ri.tolist() == [0, 1, 0, 1, 0] and cv.tolist() == [1, 4, 2, 3, 0]


# ----------------------------------------------------------------------- column sort
def test_colsort_is_stable_argsort():
    rng = np.random.default_rng(2)
    counts = rng.integers(0, 5, 500)  # many ties
    cp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ri = rng.integers(0, 100, cp[-1]).astype(np.int32)
    v = rng.standard_normal(cp[-1])
    x = rng.standard_normal(500)
    perm, sortnnz, tr, tv, sx = ob.colsort(500, cp, ri, v, x)
    ref = np.argsort(counts, kind="stable")
    assert np.array_equal(perm, ref) and np.array_equal(sortnnz, counts[ref]) and np.array_equal(sx, x[ref])
    assert np.array_equal(tr, np.concatenate([ri[cp[c]:cp[c + 1]] for c in ref]))


def test_colsort_agrees_with_reference_twin_without_ties():
    # the reference CPU twin sorts with std::sort (unstable, COO:100); on distinct keys the
    # order is fully determined, so both must agree exactly
    if not ob.ref_available(32):
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    counts = rng.permutation(np.arange(1, 301))
    cp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ri = rng.integers(0, 1000, cp[-1]).astype(np.int32)
    v = rng.standard_normal(cp[-1])
    x = rng.standard_normal(300)
    perm, sortnnz, tr, tv, sx = ob.colsort(300, cp, ri, v, x)
    rs = ob.ref_col_sort(300, cp, ri, v, x)
    assert np.array_equal(sortnnz, rs[0]) and np.array_equal(tr, rs[1]) and np.array_equal(tv, rs[2])
    assert np.array_equal(sx, rs[3])


# ----------------------------------------------------------- levels / tiles / layout
def _check_layout_invariants(t, n, nnz, W):
    T = t["tilenum"]
    assert t["tile_ptr"][T] == nnz  # COO:264
    assert np.array_equal(np.diff(t["tile_ptr"]), t["tile_len"])
    # every element of the sorted columns appears exactly once: place tmp index as value
    src = t["tmp_index"]
    assert np.array_equal(np.sort(src), np.arange(nnz))
    # element (tile t, lane l) is depth(t) of sorted column colidx+l
    P = np.concatenate([[0], np.cumsum(t["sortnnz"])])
    for u in range(T):
        for l in range(t["tile_len"][u]):
            assert src[t["tile_ptr"][u] + l] == P[t["tile_colidx"][u] + l] + t["tile_depth"][u]
    assert t["segmentoffset"][: t["segsum"]].sum() == nnz  # COO:421


@pytest.mark.parametrize("W", [32, 64])
def test_layout_matches_reference_golden_1138(W):
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    n, nnz, m = int(g["n"]), int(g["nnz"]), int(g["m"])
    t = ob.pipeline(m, n, g["csc_colptr"], g["csc_rowidx"], g["csc_val"], g["x"], W)
    assert np.array_equal(t["perm"], g["perm"]) and np.array_equal(t["sortx"], g["sortx"])
    assert t["L"] == 14 and t["tilenum"] == {32: 139, 64: 76}[W]
    keys = ["tile_ptr", "tile_len", "tile_colidx", "sortrowidx", "sortval", "sortrowindex", "countrow", "segmentoffset"]
    if W == 32:
        keys.append("tile_format")  # the reference classifies against 32 (COO:341,442)
    for k in keys:
        assert np.array_equal(t[k], g[f"w{W}_{k}"]), k
    assert t["segsum"] == int(g[f"w{W}_segsum"])


def test_layout_matches_reference_golden_symgeom():
    g = np.load(GOLDEN / "ref_layout_symgeom.npz")
    n, nnz, m = int(g["n"]), int(g["nnz"]), int(g["m"])
    for W in (32, 64):
        t = ob.pipeline(m, n, g["csc_colptr"], g["csc_rowidx"], g["csc_val"], g["x"], W)
        for k in ["tile_ptr", "tile_len", "tile_colidx", "sortrowidx", "sortval", "sortrowindex"]:
            assert np.array_equal(t[k], g[f"w{W}_{k}"]), (W, k)
        y = ob.fastload_cpu_f64(t, t["sortx"], m)
        assert np.array_equal(y, g[f"w{W}_y"])  # same serial order as FastLoad_cpu -> bitwise


@pytest.mark.parametrize("seed", range(25))
def test_layout_invariants_random(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 100))
    m = int(rng.integers(1, 80))
    W = int(rng.choice([1, 3, 32, 64, 70]))
    counts = rng.integers(0, min(m, 12) + 1, n)
    cp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    nnz = int(cp[-1])
    ri = np.concatenate([np.sort(rng.choice(m, c, replace=False)) for c in counts]).astype(np.int32) if nnz else np.zeros(0, np.int32)
    perm, sortnnz, tr, tv, sx = ob.colsort(n, cp, ri, np.zeros(nnz), np.zeros(n))
    t = ob.transform(n, nnz, sortnnz, tr, np.arange(nnz, dtype=np.float64), W)
    t["tmp_index"] = t["sortval"].astype(np.int64)
    t["sortnnz"] = sortnnz
    # closed form: tilenum = sum over levels of ceil((n-col)/W) * tall (FT:85-98)
    s = sortnnz
    steps = [(i, (s[i] - (s[i - 1] if i else 0))) for i in range(n) if s[i] != (s[i - 1] if i else 0)]
    assert t["L"] == len(steps)
    assert t["tilenum"] == sum(-(-(n - i) // W) * tall for i, tall in steps)
    # format-0 count = sum tall_k * [(n - col_k) mod W != 0] (only meaningful for W=32 semantics)
    if nnz:
        _check_layout_invariants(t, n, nnz, W)


def test_layout_brute_force_tiny():
    # every column-length pattern of n <= 4 columns with lengths <= 3 against a direct
    # construction: walk depth by depth over the suffix of columns that still have an element
    import itertools
    for n in range(1, 5):
        for counts in itertools.product(range(4), repeat=n):
            s = np.array(sorted(counts), np.int32)
            nnz = int(s.sum())
            if nnz == 0:
                continue
            W = 2
            t = ob.transform(n, nnz, s, np.arange(nnz, dtype=np.int32), np.zeros(nnz), W)
            P = np.concatenate([[0], np.cumsum(s)])
            expect = []
            depth = 0
            # levels in order of the distinct lengths; within a level chunks then depths
            distinct = sorted(set(int(v) for v in s if v > 0))
            prev = 0
            for L in distinct:
                first = int(np.searchsorted(s, L, side="left"))
                cols = list(range(first, n))
                chunks = [cols[i:i + W] for i in range(0, len(cols), W)]
                for ch in chunks:
                    for d in range(prev, L):
                        expect.append([int(P[c] + d) for c in ch])
                prev = L
            got = [t["sortrowidx"][t["tile_ptr"][u]:t["tile_ptr"][u + 1]].tolist() for u in range(t["tilenum"])]
            assert got == expect, (counts, got, expect)


def test_classification_runs():
    # rows [5,5,5,7,7,1] in one 6-wide tile -> heads at 0,3,5 with runs 3,2,1; format 2
    s = np.array([1] * 6, np.int32)
    t = ob.transform(6, 6, s, np.array([5, 5, 5, 7, 7, 1], np.int32), np.zeros(6), 6)
    assert t["sortrowindex"].tolist() == [3, 0, 0, 2, 0, 1]
    assert t["tile_format"].tolist() == [2] and t["segsum"] == 3 and t["segmentoffset"].tolist() == [3, 2, 1, 0]
    t = ob.transform(6, 6, s, np.array([4] * 6, np.int32), np.zeros(6), 6)
    assert t["tile_format"].tolist() == [1] and t["sortrowindex"].tolist() == [6, 0, 0, 0, 0, 0]
    t = ob.transform(6, 6, s, np.array([4] * 6, np.int32), np.zeros(6), 8)
    assert t["tile_format"].tolist() == [0]  # partial tile


# ------------------------------------------------------------------------------ SpMV
def test_fastload_cpu_integer_data_exact_1138():
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n = int(g["m"]), int(g["n"])
    rp, ci = g["csr_rowptr"], g["csr_colidx"]
    v = (np.arange(int(g["nnz"])) % 10).astype(np.float64)
    x = g["x"]
    y_csr = ob.spmv_csr(m, rp, ci, v, x)  # y_goldencsr, MAIN:98-108
    y_csc = ob.spmv_csc(m, n, g["csc_colptr"], g["csc_rowidx"], g["csc_val"], x)  # MAIN:110-120
    assert np.array_equal(y_csr, y_csc)  # MAIN:122-132
    for W in (32, 64):
        t = ob.pipeline(m, n, g["csc_colptr"], g["csc_rowidx"], g["csc_val"], x, W)
        y = ob.fastload_cpu_f64(t, t["sortx"], m)
        assert np.array_equal(y, y_csr)  # FC:33-45
        assert np.array_equal(y, g[f"w{W}_y"])
    dense = sp.csr_matrix((v, ci, rp), shape=(m, n)).toarray()
    assert np.array_equal(dense @ x, y_csr)


def test_kat_9x9():
    kat = json.loads((GOLDEN / "kat_9x9.json").read_text())
    A = np.array(kat["A"], np.float64)
    cp, ri, cv = csc_of(A)
    assert cp[-1] == 50
    x = np.ones(9)
    for W in (2, 32, 64):
        t = ob.pipeline(9, 9, cp, ri, cv, x, W)
        assert ob.fastload_cpu_f64(t, t["sortx"], 9).tolist() == kat["row_sums"]
    y, scale = ob.spmv_oracle(9, 9, cp, ri, cv, x)
    assert y.tolist() == kat["row_sums"]


def test_compensated_oracle_vs_fsum():
    rng = np.random.default_rng(7)
    m, n = 50, 40
    A = sp.random(m, n, density=0.3, random_state=rng, format="csc")
    A.data = rng.standard_normal(A.nnz) * 10.0 ** rng.integers(-8, 8, A.nnz)
    A.sort_indices()
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    y, scale = ob.spmv_oracle(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, x)
    Ar = A.tocsr()
    for r in range(m):
        cols = Ar.indices[Ar.indptr[r]:Ar.indptr[r + 1]]
        vals = Ar.data[Ar.indptr[r]:Ar.indptr[r + 1]]
        # exact products as (hi, lo) pairs summed exactly by fsum
        terms = []
        for a, c in zip(vals, cols):
            p = a * x[c]
            terms += [p, math.fma(a, x[c], -p)] if hasattr(math, "fma") else [p]
        exact = math.fsum(terms)
        assert abs(y[r] - exact) <= 4 * np.finfo(float).eps * abs(exact) + 1e-300
        assert scale[r] == pytest.approx(sum(abs(a * x[c]) for a, c in zip(vals, cols)), rel=1e-12)


def test_bf16_split_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    v = (rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000)).astype(np.float32)
    v[:4] = [0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9]  # ties to even
    hi, lo = ob.bf16_split(v)
    th = torch.from_numpy(v).to(torch.bfloat16)
    assert np.array_equal(hi, th.view(torch.int16).numpy().view(np.uint16))
    hif = (hi.astype(np.uint32) << 16).view(np.float32)
    tl = torch.from_numpy(v - hif).to(torch.bfloat16)
    assert np.array_equal(lo, tl.view(torch.int16).numpy().view(np.uint16))
    # hi+lo reproduces v to about 2^-16 relative
    lof = (lo.astype(np.uint32) << 16).view(np.float32)
    nz = v != 0
    rel = np.abs((hif.astype(np.float64) + lof) - v)[nz] / np.abs(v[nz])
    assert rel.max() <= 2.0 ** -16


# --------------------------------------------------------- pins of the remaining oracle functions
@pytest.mark.parametrize("a", [[], [5], [3, 0, 4, 1, 1, 0, 9], list(range(100, 0, -3))])
def test_exclusive_scan_matches_cumsum(a):
    """or_exclusive_scan (AUX:29-44) against numpy: out[0] = 0, out[i] = sum(a[:i])."""
    a = np.asarray(a, np.int32)
    want = np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.int32) if a.size else a
    assert np.array_equal(ob.exclusive_scan(a), want)
    rng = np.random.default_rng(7)
    b = rng.integers(0, 1000, 5000).astype(np.int32)
    assert np.array_equal(ob.exclusive_scan(b), np.concatenate([[0], np.cumsum(b)[:-1]]))


def _tiles_of(g, W):
    return dict(tilenum=int(g[f"w{W}_tilenum"]), tile_ptr=g[f"w{W}_tile_ptr"], tile_len=g[f"w{W}_tile_len"],
                tile_colidx=g[f"w{W}_tile_colidx"], sortrowidx=g[f"w{W}_sortrowidx"], sortval=g[f"w{W}_sortval"])


@pytest.mark.parametrize("W", [32, 64])
def test_fastload_cpu_f32_integer_data_matches_reference_y(W):
    """or_fastload_cpu_f32 on the FP32-exact integer data of src/main.cu:68-69,93-96 (values
    i % 10, x i % 10, every partial sum an integer below 2^24) equals the y the reference's own
    FastLoad_cpu produced in FP64 (the committed fixture), bit for bit."""
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m = int(g["m"])
    t = _tiles_of(g, W)
    y = ob.fastload_cpu_f32(t, t["sortval"].astype(np.float32), g["sortx"].astype(np.float32), m)
    assert np.abs(g[f"w{W}_y"]).max() < 2 ** 24
    assert np.array_equal(y.astype(np.float64), g[f"w{W}_y"])


@pytest.mark.parametrize("W", [32, 64])
def test_fastload_cpu_f32_matches_serial_float32_loop(W):
    """or_fastload_cpu_f32 on random data against a float32 numpy loop in tile order: every
    product and every sum rounded to FP32, in the order of FC:19-31 -- bit for bit."""
    g = np.load(GOLDEN / "ref_layout_symgeom.npz")
    m = int(g["m"])
    t = _tiles_of(g, W)
    rng = np.random.default_rng(W)
    v = rng.standard_normal(t["sortval"].size).astype(np.float32)
    x = rng.standard_normal(g["sortx"].size).astype(np.float32)
    want = np.zeros(m, np.float32)
    for i in range(t["tilenum"]):
        s0, s1, ln, cx = int(t["tile_ptr"][i]), int(t["tile_ptr"][i + 1]), int(t["tile_len"][i]), int(t["tile_colidx"][i])
        for j in range(s0, s1):
            r = t["sortrowidx"][j]
            want[r] = np.float32(want[r] + np.float32(v[j] * x[cx + (j - s0) % ln]))
    got = ob.fastload_cpu_f32(t, v, x, m)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ------------------------------------------------------------------------- packed value rounding
def test_pack_values_pinned_by_ieee_conversions():
    """or_pack_values rounds FP64 to m mantissa bits, nearest even: with m = 23 it is the FP32
    conversion and with m = 10 the FP16 conversion (numpy, round to nearest even) wherever those
    formats hold the value as a normal number; below 2^-126 it flushes to zero."""
    rng = np.random.default_rng(11)
    v = np.concatenate([rng.standard_normal(20000) * np.exp2(rng.integers(-40, 40, 20000)),
                        [0.0, -0.0, 1.0, -2.5, 1.5 * 2.0 ** -126, 2.0 ** -127, 3.0e38]])
    assert np.array_equal(ob.pack_values(v, 23)[:-3], v.astype(np.float32)[:-3])
    assert ob.pack_values(v, 23)[-3] == np.float32(1.5 * 2.0 ** -126) and ob.pack_values(v, 23)[-2] == 0.0
    h = rng.standard_normal(20000) * np.exp2(rng.integers(-10, 12, 20000))
    h = h[(np.abs(h) >= 2.0 ** -14) & (np.abs(h) < 65000)]  # FP16 normal numbers
    assert h.size > 15000
    assert np.array_equal(ob.pack_values(h, 10), h.astype(np.float16).astype(np.float32))


@pytest.mark.parametrize("m", [12, 13, 17])
def test_pack_values_ties_and_bounds(m):
    """Brute checks for mantissa widths the schedules use: every result has at most m mantissa
    bits, lies within half a unit of the last place, and exact halfway cases go to the even
    neighbour (including the carry into the next binade)."""
    rng = np.random.default_rng(m)
    v = rng.standard_normal(5000) * np.exp2(rng.integers(-20, 20, 5000))
    p = ob.pack_values(v, m).astype(np.float64)
    ex = np.floor(np.log2(np.abs(v)))
    ulp = np.exp2(ex - m)
    assert np.all(np.abs(p - v) <= ulp / 2 * (1 + 1e-12))
    assert np.all(np.mod(p / ulp, 1.0) == 0) or np.all(np.mod(np.abs(p) / np.exp2(np.floor(np.log2(np.abs(p))) - m), 1.0) == 0)
    one = 1.0
    tie_even = one + 2.0 ** -(m + 1)                 # halfway between 1 and 1 + 2^-m: goes down (even)
    tie_odd = one + 2.0 ** -m + 2.0 ** -(m + 1)      # halfway above an odd mantissa: goes up
    carry = 2.0 - 2.0 ** -(m + 1)                    # rounds up into the next binade
    got = ob.pack_values(np.array([tie_even, tie_odd, carry, -tie_odd]), m).astype(np.float64)
    assert got.tolist() == [1.0, 1.0 + 2.0 ** -(m - 1), 2.0, -(1.0 + 2.0 ** -(m - 1))]
