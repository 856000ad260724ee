"""CPU-only checks of the product library: it loads, exports every symbol include/*.h declares,
its host code (Matrix Market loader, generators) agrees with the oracle, and the index math of the
FAST kernel's level walk reproduces the tile layout exactly (so it never reads out of bounds)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as ob
from conftest import GOLDEN, ROOT

import fastload_amd as fl


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for mt in re.finditer(r"\b(fl_[A-Za-z0-9_]+)\s*\(", text):
            names.add(mt.group(1))
    return names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "fastload_amd" / "libfastload.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared_functions() - exported
    assert not missing, missing
    for name in declared_functions():
        assert hasattr(fl.lib, name)


def test_compare_library_exports_its_header():
    """libfastload_compare.so (measurement-only comparators) exports every flc_ symbol of
    include/fastload_compare.h, and libfastload.so does not depend on it."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "fastload_compare.h").read_text(), flags=re.S)
    names = {mt.group(1) for mt in re.finditer(r"\b(flc_[A-Za-z0-9_]+)\s*\(", text)}
    assert "flc_csc_spmv" in names and "flc_rocsparse_run" in names
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "fastload_amd" / "libfastload_compare.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert not names - exported, names - exported
    deps = subprocess.run(["readelf", "-d", str(ROOT / "fastload_amd" / "libfastload.so")], capture_output=True,
                          text=True, check=True).stdout
    assert "rocsparse" not in deps and "fastload_compare" not in deps


def test_compat_header_compiles():
    src = ROOT / "include" / "fastload_compat.hpp"
    if not src.exists():
        pytest.skip("no compat header")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_loader_matches_oracle_1138():
    a = fl.mmio_allinone(str(GOLDEN / "1138_bus.mtx"))
    b = ob.mmio_allinone(str(GOLDEN / "1138_bus.mtx"))
    assert a[:4] == b[1:5]
    for x, y in zip(a[4:], b[5:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate pattern general\n% c\n3 4 3\n1 4\n3 1\n2 2\n",
    "%%MatrixMarket matrix coordinate integer symmetric\n3 3 3\n2 1 5\n3 3 -1\n3 1 7\n",
    "%%MatrixMarket matrix coordinate complex hermitian\n2 2 2\n1 1 1.5 0\n2 1 2.5 -1\n",
    "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2\n",
    "%%MatrixMarket MATRIX Coordinate REAL General\n2 3 2\n1 3 1e-3\n2 1 -7.25\n",
    "%%MatrixMarket matrix coordinate real general\n\n2 2 1\n2 2 3\n",
])
def test_loader_variants_match_oracle(tmp_path, text):
    p = tmp_path / "a.mtx"
    p.write_text(text)
    a = fl.mmio_allinone(str(p))
    b = ob.mmio_allinone(str(p))
    assert b[0] == 0 and a[:4] == b[1:5]
    for x, y in zip(a[4:], b[5:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("text,code", [
    ("junk\n", -2), ("%%MatrixMarket matrix coordinate real general\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n", -3),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1\n", -5),
    ("%%MatrixMarket matrix array real general\n1 1\n1\n", -6),
])
def test_loader_errors_match_oracle(tmp_path, text, code):
    p = tmp_path / "b.mtx"
    p.write_text(text)
    with pytest.raises(fl.FastLoadError) as e:
        fl.mmio_allinone(str(p))
    assert e.value.rc == code == ob.mmio_allinone(str(p))[0]
    with pytest.raises(fl.FastLoadError):
        fl.mmio_allinone(str(p) + ".missing")


def _rows_ascending(cp, ri):
    for c in range(len(cp) - 1):
        seg = ri[cp[c]:cp[c + 1]]
        if len(seg) > 1 and not np.all(np.diff(seg) > 0):
            return False
    return True


def test_generators_basic():
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_SYMGEOM, 1138, 1138, 1138)
    import scipy.sparse as sp
    A = sp.csc_matrix((v, ri, cp), shape=(m, n))
    assert (A != A.T).nnz == 0  # structurally (and numerically) symmetric
    assert np.all(A.diagonal() >= 3.0)
    assert 3900 < len(ri) < 4300 and _rows_ascending(cp, ri)
    offdiag = np.diff(cp) - 1
    assert offdiag.max() <= 17 and 2.3 < offdiag.mean() < 2.9
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_STENCIL27, 5, 0, 0, (6,))
    assert m == n == 216 and len(ri) == (6 + 2 * 5) ** 3 and _rows_ascending(cp, ri)
    assert np.all((np.abs(v) >= 0.5) & (np.abs(v) <= 1.5))
    for fam, params in [(fl.GEN_ZIPF, (1.8, 500)), (fl.GEN_BANDED, (9,)), (fl.GEN_BLOCKDIAG, (40, 7)),
                        (fl.GEN_UNIFORM, (1e-2,)), (fl.GEN_ZIPF_DENSE, (2.0, 100, 0.01, 400))]:
        m, n, cp, ri, v = fl.gen_matrix(fam, 3, 1000, 800, params)
        assert cp[-1] == len(ri) and _rows_ascending(cp, ri) and ri.min() >= 0 and ri.max() < 1000
        again = fl.gen_matrix(fam, 3, 1000, 800, params)
        assert np.array_equal(again[3], ri) and np.array_equal(again[4], v)


def test_vectors():
    assert fl.gen_vector(fl.VEC_MOD10, 0, 25).tolist() == [i % 10 for i in range(25)]
    u = fl.gen_vector(fl.VEC_UNIFORM, 5, 100000)
    assert -1 <= u.min() and u.max() <= 1 and abs(u.mean()) < 0.01
    g = fl.gen_vector(fl.VEC_NORMAL, 5, 100000)
    assert abs(g.std() - 1) < 0.02


# --------------------------------------------- FAST kernel level walk, emulated on the CPU
def fast_walk(n, W, lvl_col, lvl_tall, tileoff, nnzoff, T, U=4):
    """Mirror of the wave-uniform index math in k_spmv_fast64 (fastload_amd/csrc/fl_spmv.hip)."""
    groups = (T + 63) // 64
    group_level = [int(np.searchsorted(tileoff, g * 64, side="right") - 1) for g in range(groups)]
    out = []
    for gi in range(groups):
        t0, t1 = gi * 64, min(gi * 64 + 64, T)
        k = group_level[gi]
        col, tall, noff = lvl_col[k], lvl_tall[k], nnzoff[k]
        chunks = (n - col + W - 1) // W
        local = t0 - tileoff[k]
        j, e = local // tall, local % tall
        for t in range(t0, t1, U):
            for u in range(U):
                if t + u < t1:
                    width = n - col
                    ln = width - (chunks - 1) * W if j == chunks - 1 else W
                    out.append((noff + j * tall * W + e * ln, ln, col + W * j))
                    e += 1
                    if e == tall:
                        e = 0
                        j += 1
                        if j == chunks and t + u + 1 < t1:
                            k += 1
                            col, tall, noff = lvl_col[k], lvl_tall[k], nnzoff[k]
                            chunks = (n - col + W - 1) // W
                            j = 0
    return out


@pytest.mark.parametrize("seed", range(12))
def test_fast_walk_reproduces_layout(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 3000))
    if seed % 3 == 0:
        counts = rng.zipf(1.7, n).clip(0, 400)
    elif seed % 3 == 1:
        counts = rng.integers(0, 4, n)
    else:
        counts = np.full(n, 27)
    s = np.sort(counts).astype(np.int32)
    nnz = int(s.sum())
    if nnz == 0:
        return
    t = ob.transform(n, nnz, s, np.zeros(nnz, np.int32), np.zeros(nnz), 64)
    tiles_per_level = [(-(-(n - c) // 64)) * tl for c, tl in zip(t["lvl_col"], t["lvl_tall"])]
    tileoff = np.concatenate([[0], np.cumsum(tiles_per_level)]).astype(np.int64)
    nnzoff = np.concatenate([[0], np.cumsum([(n - c) * tl for c, tl in zip(t["lvl_col"], t["lvl_tall"])])])
    walk = fast_walk(n, 64, t["lvl_col"], t["lvl_tall"], tileoff, nnzoff, t["tilenum"])
    assert len(walk) == t["tilenum"]
    ptr, ln, col = map(np.array, zip(*walk))
    assert np.array_equal(ptr, t["tile_ptr"][:-1]) and np.array_equal(ln, t["tile_len"])
    assert np.array_equal(col, t["tile_colidx"])
    assert np.all(ptr + ln <= nnz) and np.all(col + ln <= n)


@pytest.mark.parametrize("seed", range(6))
def test_tile_len_closed_form_w64(seed):
    """The batched schedules derive a tile's length from its first column alone:
    len = min(64, n - colidx) for every 64-wide FastLoad tile."""
    rng = np.random.default_rng(50 + seed)
    n = int(rng.integers(1, 5000))
    s = np.sort(rng.zipf(1.6, n).clip(0, 300)).astype(np.int32)
    nnz = int(s.sum())
    t = ob.transform(n, nnz, s, np.zeros(nnz, np.int32), np.zeros(nnz), 64)
    assert np.array_equal(t["tile_len"], np.minimum(64, n - t["tile_colidx"]))


def test_batch_assemble_block_diagonal():
    import scipy.sparse as sp
    mats = [fl.gen_matrix(fl.GEN_ZIPF, i, 30 + 17 * i, 20 + 11 * i, (2.0, 50)) for i in range(6)]
    mats.append((4, 3, np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    M, N, cp, ri, v, ro, co = fl.batch_assemble(mats)
    B = sp.block_diag([sp.csc_matrix((m[4], m[3], m[2]), shape=(m[0], m[1])) for m in mats]).tocsc()
    B.sort_indices()
    assert (M, N) == B.shape and np.array_equal(cp, B.indptr) and np.array_equal(ri, B.indices)
    assert np.array_equal(v, B.data)
    assert ro.tolist() == list(np.cumsum([0] + [m[0] for m in mats]))
    assert co.tolist() == list(np.cumsum([0] + [m[1] for m in mats]))
    with pytest.raises(fl.FastLoadError):  # row index outside its block
        fl.batch_assemble([(2, 1, np.array([0, 1], np.int32), np.array([5], np.int32), np.ones(1))])


# ------------------------------------------------------------- FastLoad_cpu (product, host)
def _golden_tiles(g, W):
    return fl.Tiles(int(g[f"w{W}_tilenum"]), int(g[f"w{W}_segsum"]), g[f"w{W}_tile_ptr"], g[f"w{W}_tile_len"],
                    g[f"w{W}_tile_colidx"], g[f"w{W}_tile_format"], g[f"w{W}_sortrowidx"], g[f"w{W}_sortval"],
                    g[f"w{W}_sortrowindex"], g[f"w{W}_countrow"], g[f"w{W}_segmentoffset"])


@pytest.mark.parametrize("W", [32, 64])
def test_fastload_cpu_serial_bitwise_vs_reference(W):
    """fl_FastLoad_cpu with one thread is the reference loop (src/FastLoad_CPU.h:19-31) with its
    roundings: on float data its y equals the reference's own FastLoad_cpu output (fixture written
    from the reference source) bit for bit."""
    g = np.load(GOLDEN / "ref_layout_symgeom.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    y, errors = fl.FastLoad_cpu(_golden_tiles(g, W), nnz, m, n, g["sortx"], threads=1)
    assert errors == 0
    assert np.array_equal(y.view(np.uint64), g[f"w{W}_y"].view(np.uint64))


@pytest.mark.parametrize("threads", [0, 2, 3, 8])
def test_fastload_cpu_threads(threads, capfd):
    """Several threads: exact on integer data (the reference's own check passes), within FP64
    round-off of the serial result on float data."""
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    y, errors = fl.FastLoad_cpu(_golden_tiles(g, 32), nnz, m, n, g["sortx"], y_golden=g["w32_y"], threads=threads)
    assert errors == 0 and np.array_equal(y, g["w32_y"])
    assert "NO PASS" not in capfd.readouterr().out
    s = np.load(GOLDEN / "ref_layout_symgeom.npz")
    ms, ns, nzs = int(s["m"]), int(s["n"]), int(s["nnz"])
    y2, _ = fl.FastLoad_cpu(_golden_tiles(s, 32), nzs, ms, ns, s["sortx"], threads=threads)
    scale = np.abs(s["w32_y"]).max()
    assert np.abs(y2 - s["w32_y"]).max() <= 1e-13 * scale


def test_fastload_cpu_golden_check_counts_mismatches(capfd):
    """src/FastLoad_CPU.h:33-45: exact comparison with y_goldencsc, mismatches counted and
    printed; the call then returns 1 (not an error)."""
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    bad = g["w32_y"].copy()
    bad[[3, 77, 500]] += 1.0
    y, errors = fl.FastLoad_cpu(_golden_tiles(g, 32), nnz, m, n, g["sortx"], y_golden=bad)
    assert errors == 3
    assert "Check NO PASS! error_count_check_cpu= 3" in capfd.readouterr().out


def test_fastload_cpu_f32_matches_oracle():
    g = np.load(GOLDEN / "ref_layout_symgeom.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    t = _golden_tiles(g, 64)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(nnz).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    want = ob.fastload_cpu_f32(dict(tilenum=t.tilenum, tile_ptr=t.tile_ptr, tile_len=t.tile_len,
                                    tile_colidx=t.tile_colidx, sortrowidx=t.sortrowidx), v, x, m)
    got = fl.FastLoad_cpu_f32(t, v, nnz, m, n, x, threads=1)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    got4 = fl.FastLoad_cpu_f32(t, v, nnz, m, n, x, threads=4)
    assert np.abs(got4 - want).max() <= 1e-5 * np.abs(want).max()


def test_fastload_cpu_rejects_bad_layouts():
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    t = _golden_tiles(g, 32)
    bad = fl.Tiles(**{**t.__dict__, "tile_len": t.tile_len + 1})  # element counts no longer match lengths
    with pytest.raises(fl.FastLoadError) as e:
        fl.FastLoad_cpu(bad, nnz, m, n, g["sortx"])
    assert e.value.rc == -8
    with pytest.raises(fl.FastLoadError) as e:  # columns beyond colA
        fl.FastLoad_cpu(t, nnz, m, n - 5, g["sortx"])
    assert e.value.rc == -8
    rows = t.sortrowidx.copy()
    rows[10] = m + 3
    with pytest.raises(fl.FastLoadError) as e:
        fl.FastLoad_cpu(fl.Tiles(**{**t.__dict__, "sortrowidx": rows}), nnz, m, n, g["sortx"], threads=2)
    assert e.value.rc == -5


# ------------------------------------------------------------------- input validation (host)
def test_batch_assemble_rejects_bad_colptr():
    good = (3, 3, np.array([0, 1, 2, 3], np.int32), np.array([0, 1, 2], np.int32), np.ones(3))
    for cp in ([0, 10, 0, 3], [1, 1, 2, 3], [0, 2, 1, 3]):
        bad = (3, 3, np.array(cp, np.int32), np.array([0, 1, 2], np.int32), np.ones(3))
        with pytest.raises(fl.FastLoadError) as e:
            fl.batch_assemble([good, bad])
        assert e.value.rc == -8, cp


def _fake_matrix(m, n):
    M = object.__new__(fl.FastLoadMatrix)  # no device: only the host-side checks run
    M.m, M.n, M.handle = m, n, None
    return M


def _buf(count, dtype):
    from types import SimpleNamespace
    return SimpleNamespace(count=count, dtype=np.dtype(dtype), ptr=None, nbytes=count * np.dtype(dtype).itemsize)


@pytest.mark.parametrize("call", [
    lambda M: M.spmv(_buf(9, np.float32), _buf(7, np.float32), fl.KERNEL_PB, fl.VAL_F32),   # x too short
    lambda M: M.spmv(_buf(10, np.float32), _buf(6, np.float32), fl.KERNEL_PB, fl.VAL_F32),  # y too short
    lambda M: M.spmv(_buf(10, np.float64), _buf(7, np.float32), fl.KERNEL_PB, fl.VAL_F32),  # x dtype
    lambda M: M.spmv(_buf(10, np.float32), _buf(7, np.float32), fl.KERNEL_PB, fl.VAL_BF16X2),  # y must be f64
    lambda M: M.spmv(_buf(10, np.float32), _buf(7, np.float64), fl.KERNEL_FAST, fl.VAL_BF16X2,
                     work=_buf(13, np.float32)),  # work needs 2m floats
    lambda M: M.spmm(_buf(10 * 8 - 1, np.float32), _buf(7 * 8, np.float32), 8),  # X needs n*k
    lambda M: M.spmm(_buf(10 * 8, np.float32), _buf(7 * 4, np.float32), 8),      # Y sized for k=4
    lambda M: M.spmm(_buf(10, np.float64), _buf(7, np.float64), 1, fl.VAL_BF16),  # kind
    lambda M: M.permute_x(_buf(10, np.float64), _buf(10, np.float32)),
    lambda M: M.permute_x(_buf(9, np.float32), _buf(10, np.float32)),
])
def test_wrappers_refuse_bad_buffers(call):
    """FastLoadMatrix.spmv/spmm/permute_x compare buffer sizes and element types with the matrix
    and the value kind before any pointer reaches the library."""
    with pytest.raises(fl.FastLoadError) as e:
        call(_fake_matrix(7, 10))
    assert e.value.rc == -8


# ------------------------------------------------------------- slide_matrix file (host)
@pytest.mark.parametrize("name,W", [("ref_layout_1138_bus", 32), ("ref_layout_1138_bus", 64), ("ref_layout_symgeom", 32),
                                    ("ref_layout_symgeom", 64)])
def test_slide_file_round_trip(tmp_path, name, W):
    """fl_slide_save -> fl_slide_load gives back every slide_matrix array bit for bit, and the loaded
    layout multiplies exactly as the original (FastLoad_cpu, serial)."""
    g = np.load(GOLDEN / f"{name}.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    t = _golden_tiles(g, W)
    p = tmp_path / "a.flmx"
    fl.slide_save(p, t, m, n, nnz, 14, perm=g["perm"])
    t2, m2, n2, nnz2, h2, perm2 = fl.slide_load(p)
    assert (m2, n2, nnz2, h2) == (m, n, nnz, 14)
    assert np.array_equal(perm2, g["perm"])
    for k in ("tile_ptr", "tile_len", "tile_colidx", "tile_format", "sortrowidx", "sortrowindex", "countrow",
              "segmentoffset"):
        assert np.array_equal(getattr(t2, k), getattr(t, k)), k
    assert np.array_equal(t2.sortval.view(np.uint64), t.sortval.view(np.uint64))
    assert (t2.tilenum, t2.segsum) == (t.tilenum, t.segsum)
    y1, _ = fl.FastLoad_cpu(t, nnz, m, n, g["sortx"])
    y2, _ = fl.FastLoad_cpu(t2, nnz, m, n, g["x"][perm2])
    assert np.array_equal(y1.view(np.uint64), y2.view(np.uint64))


def test_slide_file_format_and_damage(tmp_path):
    """The header starts with "FLMATRIX", version 1 and the byte-order mark 0x01020304; a file whose
    magic, version, length or section contents were changed is refused, a missing file gives
    FL_ERR_OPEN, and a file without a column order loads with perm None."""
    import struct
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    t = _golden_tiles(g, 32)
    p = tmp_path / "b.flmx"
    fl.slide_save(p, t, m, n, nnz, 14)
    raw = bytearray(p.read_bytes())
    assert raw[:8] == b"FLMATRIX" and struct.unpack_from("<II", raw, 8) == (1, 0x01020304)
    assert struct.unpack_from("<iii", raw, 16) == (m, n, nnz)
    assert fl.slide_load(p)[5] is None

    def refused(data, rc=-8):
        q = tmp_path / "bad.flmx"
        q.write_bytes(bytes(data))
        with pytest.raises(fl.FastLoadError) as e:
            fl.slide_load(q)
        assert e.value.rc == rc

    bad = bytearray(raw); bad[0] ^= 1; refused(bad)                     # magic
    bad = bytearray(raw); bad[8] = 2; refused(bad)                      # version
    refused(raw[: len(raw) - 100])                                      # truncated: a section leaves the file
    bad = bytearray(raw); bad[-5] ^= 0x40; refused(bad)                 # last section's data: hash mismatch
    with pytest.raises(fl.FastLoadError) as e:
        fl.slide_load(tmp_path / "missing.flmx")
    assert e.value.rc == -1


def test_slide_file_empty_matrix(tmp_path):
    t = fl.Tiles(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                 np.zeros(0, np.int32), np.zeros(0), np.zeros(0, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32))
    p = tmp_path / "e.flmx"
    fl.slide_save(p, t, 3, 4, 0, 0, perm=np.arange(4, dtype=np.int32))
    t2, m, n, nnz, h, perm = fl.slide_load(p)
    assert (m, n, nnz, t2.tilenum) == (3, 4, 0, 0) and perm.tolist() == [0, 1, 2, 3]


def test_slide_file_hashes_and_tamper_detection(tmp_path):
    """The test-side restatement of the section hash agrees with the library's for every section;
    a section changed and re-hashed still loads (the hash guards accidents; the host API validates
    layouts when they are used, as FastLoad_cpu's refusal shows)."""
    from flmx_util import section_hash, sections, tamper
    g = np.load(GOLDEN / "ref_layout_1138_bus.npz")
    m, n, nnz = int(g["m"]), int(g["n"]), int(g["nnz"])
    p = tmp_path / "h.flmx"
    fl.slide_save(p, _golden_tiles(g, 32), m, n, nnz, 14, perm=g["perm"])
    raw = p.read_bytes()
    secs = sections(raw)
    assert set(secs) == {1, 3, 4, 5, 6, 7, 8, 9, 10, 11}
    for tag, (_, elem, count, off, h) in secs.items():
        assert section_hash(raw[off:off + elem * count]) == h, tag
    q = tmp_path / "t.flmx"
    q.write_bytes(bytes(tamper(raw, 4, 0, 31, "i")))  # tile 0's length no longer matches its elements
    t2, *_ = fl.slide_load(q)
    with pytest.raises(fl.FastLoadError) as e:
        fl.FastLoad_cpu(t2, nnz, m, n, g["sortx"])
    assert e.value.rc == -8


def test_stream_policy_values():
    """fl_set_stream_policy accepts the three policies and refuses anything else (host only: it
    sets a process-wide switch and enqueues nothing)."""
    try:
        for p in (fl.STREAM_NEVER, fl.STREAM_ALWAYS, fl.STREAM_AUTO):
            assert fl.lib.fl_set_stream_policy(p) == 0
        for bad in (2, -2, 100):
            assert fl.lib.fl_set_stream_policy(bad) == -8
            assert b"fl_set_stream_policy" in fl.lib.fl_last_error()
        with pytest.raises(fl.FastLoadError):
            fl.set_stream_policy(7)
    finally:
        fl.set_stream_policy(fl.STREAM_AUTO)


@pytest.mark.parametrize("g", [1, 2, 5])
def test_laplacian27_is_the_spd_kronecker_form(g):
    """The CG workload: 27 I - T (x) T (x) T with T = tridiag(1, 1, 1), symmetric, eigenvalues in
    (0, 54) (those of T lie in (-1, 3), so the product's lie in (-27, 27))."""
    import scipy.sparse as sp
    from fastload_amd import workloads as wl
    m, n, cp, ri, v, b = wl.laplacian27(g)
    assert m == n == g ** 3 and len(b) == n and cp[n] == (3 * g - 2) ** 3
    A = sp.csc_matrix((v, ri, cp), shape=(m, n)).toarray()
    T = np.eye(g) + np.eye(g, k=1) + np.eye(g, k=-1)
    assert np.array_equal(A, 27 * np.eye(n) - np.kron(np.kron(T, T), T))
    ev = np.linalg.eigvalsh(A)
    assert ev[0] > 0 and ev[-1] < 54
