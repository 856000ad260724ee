"""Writes the golden fixtures under tests/golden/ (run in the build container, where the reference
sources exist; the GPU hosts only read the committed outputs).

  1138_bus.mtx            the reference's only input file (Data/1138_bus.mtx), copied as data
  ref_layout_1138_bus.npz tile layout + classification produced by the reference's own CPU
                          twins (oracle/_ref: src/coo2csc.h formattransation/tile_classification,
                          compiled from the reference headers) at slidesize 32 and 64, plus
                          FastLoad_cpu's y (src/FastLoad_CPU.h) on the integer data of
                          src/main.cu:68-69,93-96
  ref_layout_symgeom.npz  the same for the seeded 1138x1138 structurally symmetric twin
  kat_9x9.json            the hand-written 9x9 matrix of spmv.cu:140-150 and its row sums

The stable column sort that feeds formattransation is numpy argsort(kind="stable"), the
semantics of thrust::sort_by_key on int keys (src/ColSort_GPU.h:101).
"""
from __future__ import annotations

import json
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT))

import oracle_bind as ob  # noqa: E402

REF_MTX = Path("/root/reference/Data/1138_bus.mtx")


def stable_sorted_columns(n, colptr, rowidx, val, x):
    counts = np.diff(colptr).astype(np.int32)
    perm = np.argsort(counts, kind="stable").astype(np.int32)
    tmp_row = np.concatenate([rowidx[colptr[c]:colptr[c + 1]] for c in perm]) if len(rowidx) else rowidx
    tmp_val = np.concatenate([val[colptr[c]:colptr[c + 1]] for c in perm]) if len(val) else val
    return perm, counts[perm], tmp_row.astype(np.int32), tmp_val.astype(np.float64), x[perm]


def ref_fixture(m, n, colptr, rowidx, val, x):
    out = {}
    perm, sortnnz, tmp_row, tmp_val, sortx = stable_sorted_columns(n, colptr, rowidx, val, x)
    nnz = int(colptr[n])
    out.update(perm=perm, sortnnz=sortnnz, sortx=sortx, tmp_row=tmp_row, tmp_val=tmp_val)
    for W in (32, 64):
        A, r = ob.ref_transform_classify(W, m, n, nnz, sortnnz, tmp_row, tmp_val)
        y = ob.ref_fastload_cpu(A, nnz, m, n, sortx, W)
        ob.ref_destroy(A, W)
        for k, v in r.items():
            out[f"w{W}_{k}"] = np.asarray(v)
        out[f"w{W}_y"] = y
    return out


def main():
    if not ob.ref_available(32) or not ob.ref_available(64):
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` first")
    if REF_MTX.exists():
        shutil.copyfile(REF_MTX, HERE / "1138_bus.mtx")
    mtx = HERE / "1138_bus.mtx"

    # 1138_bus through the reference loader + csr2csc, values/x = i % 10 (src/main.cu:68-69,93-96)
    rc, m, n, nnz, sym, rp, ci, v = ob.ref_mmio_allinone(str(mtx))
    assert rc == 0
    v = (np.arange(nnz) % 10).astype(np.float64)
    colptr, rowidx, cval, npc = ob.ref_csr2csc(m, n, rp, ci, v)
    x = (np.arange(n) % 10).astype(np.float64)
    fx = ref_fixture(m, n, colptr, rowidx, cval, x)
    fx.update(m=m, n=n, nnz=nnz, csr_rowptr=rp, csr_colidx=ci, csc_colptr=colptr, csc_rowidx=rowidx, csc_val=cval, x=x)
    np.savez_compressed(HERE / "ref_layout_1138_bus.npz", **fx)

    # seeded structurally symmetric twin (float data): layout only depends on the pattern
    import fastload_amd as fl
    m2, n2, cp2, ri2, v2 = fl.gen_matrix(fl.GEN_SYMGEOM, 1138, 1138, 1138)
    x2 = fl.gen_vector(fl.VEC_UNIFORM, 1139, n2)
    fx2 = ref_fixture(m2, n2, cp2, ri2, v2, x2)
    fx2.update(m=m2, n=n2, nnz=int(cp2[n2]), csc_colptr=cp2, csc_rowidx=ri2, csc_val=v2, x=x2)
    np.savez_compressed(HERE / "ref_layout_symgeom.npz", **fx2)

    # 9x9 known-answer test (spmv.cu:140-150), x = 1: row sums
    A = [[5, 0, 0, 0, 0, 0, 1, 0, 2], [3, 4, 5, 0, 0, 0, 6, 0, 7], [0, 8, 9, 0, 0, 0, 5, 0, 1],
         [2, 0, 3, 4, 5, 6, 0, 0, 0], [7, 0, 8, 9, 5, 1, 0, 0, 0], [2, 0, 3, 4, 5, 6, 0, 0, 0],
         [7, 0, 8, 9, 5, 1, 0, 0, 0], [2, 3, 4, 5, 6, 7, 8, 9, 5], [1, 2, 3, 4, 5, 6, 7, 8, 9]]
    (HERE / "kat_9x9.json").write_text(json.dumps({"A": A, "x": [1] * 9,
                                                   "row_sums": [8, 25, 23, 20, 30, 20, 30, 49, 45]}, indent=1))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
