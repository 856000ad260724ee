"""FastLoad SpMV benchmark on one MI355X.

Headline workload (`value`): the power-law configuration -- 4,194,304 x 4,194,304 CSC, column
nnz Zipf (exponent 1.8423 so the mean is 16 with max 50,000; ~67M nnz), rows uniform without
repeats, values and x N(0,1), seed 42. A step is one y = A*x with the matrix and x resident in
HBM: y cleared by a fill kernel then one accumulating SpMV (drop-in kernels), or one
fl_spmv_assign call (batched schedules: the propagation-blocking second phase writes every row
of y; WIN/WINR clear y and store the rows a single batch owns).

Variants (schedules x value storage):
  plain_*  the reference per-tile algorithm on the reference 32-wide layout (drop-in)
  fast_*   64-wide tiles, level-table geometry, one global atomic per segment
  win_*    batches of 256 tiles, rows compacted to LDS slots, FP64 LDS sums, one flush per row
           (a store when no other batch touches the row, else one atomic)
  winr*    the same kernel on tiles grouped by row window (most rows owned by one batch)
  pb_*     two-phase propagation blocking over 16,384-row bands, FP64 LDS sums, no global atomics
  row_*    entries re-ordered by row, four lanes per row, no atomics
  hyb_*    the min(n, 32768) longest columns (19,456 in FP64) row by row with their x staged in
           LDS, the other columns through PB on a sub-matrix
  learned_* the schedule the learned selector picks for the matrix: features and MLP inference
           on the GPU (committed trained weights) rank the 13 schedules, those of probability
           >= 2 % (at most 5) are timed on the matrix and the fastest is used; in each value
           storage it supports
  *_f32 / *_f64 / *_bf16 / *_bf16x2: value storage (split hi+lo BF16 recombined in FP64)
`value` = GB/s-equivalent (algorithmic bytes with FP32 values / step time) of the fastest
learned_* variant whose error vs the FP64 path is <= 1e-4; the plain FP32 drop-in is reported
beside it. The stencil configuration (128^3 27-point, 55.7M nnz) is timed too and reported
under "stencil" with the error of every value storage. "cg": conjugate gradients in FP64 on the
resident 27-point SPD matrix on 128^3 (iterations to 1e-8, ms and GB/s per iteration);
"cache_file": saving and loading that matrix's layout and schedule against preparing them.

Algorithmic bytes per SpMV: B = nnz*(s_val + 4) + n*s_x + m*s_y.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_alg(nnz, n, m, s_val, s_x, s_y):
    return nnz * (s_val + 4) + n * s_x + m * s_y


VARIANTS = {
    # name: (W, kernel, kind, s_val, batch_tiles, band_rows); kernel -1 = the row-ordered kernel
    "plain_f32": (32, 0, 1, 4, 0, 0),
    "plain_f64": (32, 0, 0, 8, 0, 0),
    "fast_f32": (64, 1, 1, 4, 0, 0),
    "win_f32": (64, 2, 1, 4, 256, 0),
    "win_bf16x2": (64, 2, 3, 4, 256, 0),
    "win_bf16": (64, 2, 2, 2, 256, 0),
    "win_f64": (64, 2, 0, 8, 256, 0),
    "pb_f32": (64, 3, 1, 4, 256, 16384),
    "pb_bf16x2": (64, 3, 3, 4, 256, 16384),
    "pb_bf16": (64, 3, 2, 2, 256, 16384),
    "pb_f64": (64, 3, 0, 8, 256, 16384),
    "row_f32": (64, -1, 1, 4, 0, 0),
    "winr64_f32": (64, 4, 1, 4, 64, 0),
    # schedule-parameter sweep points
    "win128_f32": (64, 2, 1, 4, 128, 0),
    "win64_f32": (64, 2, 1, 4, 64, 0),
    "win128_bf16x2": (64, 2, 3, 4, 128, 0),
    "win128_bf16": (64, 2, 2, 2, 128, 0),
    "win128_f64": (64, 2, 0, 8, 128, 0),
    "pb128_f32": (64, 3, 1, 4, 128, 16384),
    "pb8k_f32": (64, 3, 1, 4, 256, 8192),
    "pb4k_f32": (64, 3, 1, 4, 256, 4096),
    "pb128_8k_f32": (64, 3, 1, 4, 128, 8192),
    # HYB (kernel 5): heaviest 32768 columns from LDS (FP32 values), the rest PB 128 x 16K; in
    # FP64 the heaviest 19456 (x takes 8 bytes in LDS)
    "hyb_f32": (64, 5, 1, 4, 128, 16384),
    "hyb_packed": (64, 5, 4, 4, 128, 16384),
    "hyb_f64": (64, 5, 0, 8, 128, 16384),
}
# value storage -> (name, value bytes beside a 4-byte row index); packed words hold the value and
# the schedule's slot in 4 bytes, so they count 0 + 4
KIND_SVAL = {0: ("f64", 8), 1: ("f32", 4), 2: ("bf16", 2), 3: ("bf16x2", 4), 4: ("packed", 0)}


def learned_variants(sched, names=("f32", "bf16x2", "bf16", "packed")):
    """Value-storage variants of the schedule the selector chose: FP32, split hi+lo BF16, BF16 and
    packed value + slot words (those of `names` that the schedule supports without a work
    buffer)."""
    kinds = ([1, 3, 2, 4] if sched.kernel in (2, 3, 4) else [1, 4] if sched.kernel == 5
             else ([1, 2] if sched.kernel >= 0 else [1]))
    # HYB keeps FP32 values for its heavy part (most entries): 4 value bytes in both kinds
    return {f"learned_{KIND_SVAL[k][0]}": (sched.width, sched.kernel, k, 4 if sched.kernel == 5 else KIND_SVAL[k][1],
                                           sched.batch_tiles, sched.band_rows) for k in kinds if KIND_SVAL[k][0] in names}


def profiled_traffic(schedule: str, kind: str, workload: str):
    """HBM bytes per SpMV from a committed rocprofv3 counter summary of this workload, schedule
    and value storage (profiles/traffic_*.md, line `traffic_bytes_per_spmv: <workload>
    <schedule> <kind> <bytes>`), or None when no such profile exists."""
    import glob
    for f in sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_*.md"))):
        for line in open(f):
            parts = line.strip().strip("`").split()
            if len(parts) == 5 and parts[0] == "traffic_bytes_per_spmv:" and parts[1:4] == [workload, schedule, kind]:
                return int(parts[4])
    return None


def time_variant(fl, run, dy, steps, warmup, zero=True, flush=None):
    """Step = (y = 0) + run(). The step time comes from a loop of `steps` steps bracketed by two
    events and nothing else; a second loop of `steps` steps records events around every run()
    (the SpMV launches) for the kernel times. With `flush` (a device buffer larger than the
    256 MB Infinity Cache), a fill kernel over it runs before every rep of the second loop,
    outside the run() events, so no rep finds the previous rep's data in the cache; only the
    per-run times are meaningful then, and the step loop is skipped (step time None)."""
    from fastload_amd import Event
    for _ in range(warmup):
        if zero:
            fl.zero_device(dy)
        run()
    fl.synchronize()
    step_ms = None
    if flush is None:
        t0, t1 = Event(), Event()
        t0.record()
        for _ in range(steps):
            if zero:
                fl.zero_device(dy)
            run()
        t1.record()
        fl.synchronize()
        step_ms = t0.elapsed_ms(t1) / steps
    ev = [(Event(), Event()) for _ in range(steps)]
    for a, b in ev:
        if flush is not None:
            fl.zero_device(flush)
        if zero:
            fl.zero_device(dy)
        a.record()
        run()
        b.record()
    fl.synchronize()
    kern = np.array([a.elapsed_ms(b) for a, b in ev])
    return step_ms, float(kern.mean()), float(np.median(kern))


def select_schedule(fl, m, n, dcp, dri, dv, top=None, force=None):
    """The learned selector on the resident matrix: features and MLP inference on the GPU rank the
    13 schedules; the schedules the network gives at least 2 % probability (at most 5; or the
    `top` best-ranked), and HYB when the 32768 longest columns hold at least 35 % of the entries,
    are timed on the matrix (model-guided tuning) and the fastest is used.
    Both the network's first choice and the tuned one are reported. `force` (a schedule name)
    skips the selection, for profiling runs whose timings the profiler skews."""
    from fastload_amd import selector as sel
    if force:
        c = next(i for i, sc in enumerate(sel.SCHEDULES) if sc.name == force)
        return sel.SCHEDULES[c], dict(cls=c, schedule=force, forced=True)
    s = sel.Selector()
    fl.synchronize()
    t0 = time.perf_counter()
    c1, p, f = s.choose(m, n, dcp, dri)
    ms1 = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    c, times, _ = s.choose_refined(m, n, dcp, dri, dv, top=top)
    ms = (time.perf_counter() - t0) * 1e3
    return sel.SCHEDULES[c], dict(cls=c, schedule=sel.SCHEDULES[c].name,
                                      network_choice=sel.SCHEDULES[c1].name, network_prob=round(p, 4),
                                      select_ms=round(ms1, 3), tuned_candidates=len(times),
                                      tuned_ms={sel.SCHEDULES[k].name: (round(v, 4) if v is not None else None)
                                                for k, v in times.items()},
                                      tuning_s=round(ms / 1e3, 3), trained=s.mlp is not None)


def run_workload(fl, name, m, n, dcp, dri, dv, x, variants, steps, warmup, flush=None):
    """Times `variants` ({name: (W, kernel, kind, s_val, bt, R)}) on the resident CSC matrix; with
    `flush`, every variant is timed a second time with the Infinity Cache flushed before each rep."""
    from fastload_amd import DeviceArray
    nnz = dri.count
    mats = {}
    for W in sorted({t[0] for t in variants.values()} | {32}):
        t0 = time.time()
        mats[W] = fl.FastLoadMatrix(m, n, dcp, dri, dv, width=W)
        fl.synchronize()
        log(f"[{name}] preprocessed W={W}: levels={mats[W].levels} tiles={mats[W].tilenum} in {time.time() - t0:.2f}s")
    dx64 = DeviceArray.from_host(x)
    dx32 = DeviceArray.from_host(x.astype(np.float32))
    sx = {}
    for W, M in mats.items():
        sx[(W, 8)] = DeviceArray(n, np.float64)
        M.permute_x(dx64, sx[(W, 8)])
        sx[(W, 4)] = DeviceArray(n, np.float32)
        M.permute_x(dx32, sx[(W, 4)])
    y64 = DeviceArray(m, np.float64)
    y32 = DeviceArray(m, np.float32)
    # FP64 reference result: the plain FP64 path (parity-tested against the FP64 oracle)
    y64.zero()
    mats[32].spmv(sx[(32, 8)], y64, fl.KERNEL_DROPIN, fl.VAL_F64)
    y_ref = y64.to_host()
    ref_norm = float(np.linalg.norm(y_ref))
    res = {}
    for vname, (W, kernel, kind, s_val, bt, R) in variants.items():
        M = mats[W]
        t0 = time.time()
        s_x = 8 if kind == fl.VAL_F64 else 4
        if kernel < 0:  # row-ordered kernel: x in the original order, y written (not accumulated)
            M.prepare_rows()
            ydev = y32

            def run(M=M):
                M.spmm(dx32, y32, 1, fl.VAL_F32)
        else:
            if kind != fl.VAL_F64:
                M.prepare(kind)
            if kernel in (fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB):
                M.prepare_schedule(kernel, bt, max(R, 1024))
            elif kernel == fl.KERNEL_HYB:
                M.prepare_hyb(min(n, 19456) if kind == fl.VAL_F64 else None, bt, R)
            ydev = y64 if kind in (fl.VAL_F64, fl.VAL_BF16X2) else y32
            assign = kernel in (fl.KERNEL_PB, fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_HYB)

            def run(M=M, kernel=kernel, kind=kind, dsx=sx[(W, s_x)], ydev=ydev, assign=assign):
                # y = A x: fl_spmv_assign for the batched schedules (PB and HYB write every row in
                # PB's second phase; WIN/WINR clear y with a fill kernel inside the call, then store
                # the rows one batch owns); the drop-in kernels accumulate into a y cleared by a
                # fill kernel outside the kernel events
                M.spmv(dsx, ydev, kernel, kind, assign=assign)
        fl.synchronize()
        prep_s = time.time() - t0
        zero = kernel >= 0 and not assign
        step_ms, kern_ms, kern_med = time_variant(fl, run, ydev, steps, warmup, zero=zero)
        flushed = None
        if flush is not None:
            _, flushed, _ = time_variant(fl, run, ydev, steps, 1, zero=zero, flush=flush)
        y = ydev.to_host().astype(np.float64)
        err = float(np.linalg.norm(y - y_ref) / ref_norm) if ref_norm else 0.0
        s_y = 8 if ydev is y64 else 4
        B = b_alg(nnz, n, m, s_val, s_x, s_y)
        B32 = b_alg(nnz, n, m, 4, 4, 4)
        res[vname] = dict(step_ms=round(step_ms, 5), kernel_ms=round(kern_ms, 5), kernel_median_ms=round(kern_med, 5),
                          gbs=round(B / step_ms / 1e6, 2), kernel_gbs=round(B / kern_ms / 1e6, 2),
                          gbs_equiv_f32=round(B32 / step_ms / 1e6, 2), kernel_gbs_equiv_f32=round(B32 / kern_ms / 1e6, 2),
                          bytes_alg=B, rel_err_l2=err, gflops=round(2 * nnz / step_ms / 1e6, 2),
                          schedule_prep_s=round(prep_s, 3))
        if flushed is not None:
            res[vname].update(kernel_ms_flushed=round(flushed, 5), kernel_gbs_flushed=round(B / flushed / 1e6, 2),
                              kernel_gbs_equiv_f32_flushed=round(B32 / flushed / 1e6, 2))
        if kind == fl.VAL_PACKED:
            mb, eb, ib = M.pack_info(kernel)
            res[vname].update(pack_mantissa_bits=mb, pack_exponent_bits=eb, pack_slot_bits=ib)
        log(f"[{name}] {vname:14s} step {step_ms:.4f} ms kernel {kern_ms:.4f} ms  {res[vname]['gbs']:7.1f} GB/s "
            f"(f32-equiv {res[vname]['gbs_equiv_f32']:7.1f})  err {err:.2e}"
            + (f"  flushed kernel {flushed:.4f} ms" if flushed is not None else ""))
    return nnz, mats, res, y_ref


def run_comparators(fl, name, m, n, dcp, dri, dv, x, y_ref, steps, warmup, flush=None):
    """Measurement-only yardsticks on the same resident matrix: the reference's naive CSC kernel
    (src/CSCSpMV.h:4-27: a lane per column, an atomic per entry) and rocSPARSE's adaptive CSR SpMV
    (the vendor library, in place of the reference's cuSPARSE call, src/cusparse_cuda.h), FP32 and
    FP64. Neither is a FastLoad schedule; they are timed like the variants (events around each
    launch; y cleared by a fill kernel first for the accumulating naive kernel)."""
    from fastload_amd import DeviceArray
    from fastload_amd import compare as cmpl
    nnz = dri.count
    cp, ri, v = dcp.to_host(), dri.to_host(), dv.to_host()
    # CSR of A = CSC of A^T: the library's own transposition, treating A's columns as rows
    cv, cr, ccp, _, _ = fl.csr_to_csc(n, m, v, cp, ri)
    del cp, ri, v
    ref_norm = float(np.linalg.norm(y_ref))
    res = {}
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        s = np.dtype(dt).itemsize
        dval = dv if dt == np.float64 else DeviceArray.from_host(dv.to_host().astype(np.float32))
        dx = DeviceArray.from_host(x.astype(dt))
        dy = DeviceArray(m, dt)
        B = b_alg(nnz, n, m, s, s, s)
        R = cmpl.RocsparseCSR(m, n, DeviceArray.from_host(ccp), DeviceArray.from_host(cr),
                              DeviceArray.from_host(cv.astype(dt)))
        for cname, run, zero in ((f"naive_csc_{tag}", lambda: cmpl.csc_naive_device(m, n, dcp, dri, dval, dx, dy), True),
                                 (f"rocsparse_csr_{tag}", lambda: R.run(dx, dy), False)):
            step_ms, kern_ms, kern_med = time_variant(fl, run, dy, steps, warmup, zero=zero)
            flushed = None
            if flush is not None:
                _, flushed, _ = time_variant(fl, run, dy, steps, 1, zero=zero, flush=flush)
            err = float(np.linalg.norm(dy.to_host().astype(np.float64) - y_ref) / ref_norm) if ref_norm else 0.0
            res[cname] = dict(step_ms=round(step_ms, 5), kernel_ms=round(kern_ms, 5),
                              kernel_median_ms=round(kern_med, 5), kernel_gbs=round(B / kern_ms / 1e6, 2),
                              bytes_alg=B, rel_err_l2=err)
            if flushed is not None:
                res[cname].update(kernel_ms_flushed=round(flushed, 5), kernel_gbs_flushed=round(B / flushed / 1e6, 2))
            log(f"[{name}] {cname:18s} kernel {kern_ms:.4f} ms  {res[cname]['kernel_gbs']:7.1f} GB/s  err {err:.2e}"
                + (f"  flushed {flushed:.4f} ms" if flushed is not None else ""))
        del R
    return res


def run_batched(fl, count, seed, steps, warmup, sequential=True, ks=(1, 8, 32)):
    """Batched configuration: `count` square CSC matrices, n log-uniform in [1e3, 2e5], column nnz
    Zipf(2.0) truncated at n, FP32, k = 1, 8, 32 right-hand sides. The batch runs as one
    block-diagonal matrix (k = 1: the schedule the learned selector picks for it; k > 1: the
    row-ordered multi-RHS kernel) captured in a hipGraph. Baselines, each `count` sequential
    per-matrix launches (eager):
      sequential_plain   the plain FP32 drop-in per matrix -- for k > 1 once per right-hand side
                         (the reference kernel multiplies one vector per launch): count * k launches
      sequential_same    the batch's own schedule run on each matrix alone (k = 1: the learned
                         schedule; k > 1: the row kernel), so the batching gain is separate from
                         the schedule gain"""
    from fastload_amd import DeviceArray, Event, Graph, Stream
    from fastload_amd import workloads as wl
    rng = np.random.default_rng(seed)
    t0 = time.time()
    mats = wl.batch(count, seed)
    sizes = np.array([m for (m, *_r) in mats])
    M_, N_, cp, ri, v, ro, co = fl.batch_assemble(mats)
    nnz = int(cp[-1])
    log(f"[batched] {count} matrices, rows {sizes.min()}..{sizes.max()}, sum rows {M_}, nnz {nnz} "
        f"generated+assembled in {time.time() - t0:.1f}s")
    s = Stream()
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    sched, chosen = select_schedule(fl, M_, N_, dcp, dri, dv)
    del dcp, dri, dv
    log(f"[batched] selector chose {chosen} for k = 1")
    B = fl.FastLoadMatrix(M_, N_, cp, ri, v, width=64)
    B.prepare_rows()
    B1 = B if sched.width == 64 else fl.FastLoadMatrix(M_, N_, cp, ri, v, width=sched.width)
    def prep_sched(Mx):
        if sched.kernel >= 0:
            Mx.prepare(fl.VAL_F32)
        if sched.kernel in (fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB):
            Mx.prepare_schedule(sched.kernel, sched.batch_tiles, max(sched.band_rows, 1024))
        elif sched.kernel == fl.KERNEL_HYB:
            Mx.prepare_hyb(None, sched.batch_tiles, sched.band_rows)
    prep_sched(B1)
    res = {"matrices": count, "rows": int(M_), "nnz": nnz, "learned_schedule_k1": chosen}

    def timed(fn, reps, warm=None):
        for _ in range(warmup if warm is None else warm):
            fn()
        s.synchronize()
        a, b = Event(), Event()
        a.record(s)
        for _ in range(reps):
            fn()
        b.record(s)
        s.synchronize()
        return a.elapsed_ms(b) / reps

    # per-matrix objects for the sequential baselines: the 32-wide drop-in layout (plain) and the
    # learned schedule's layout (same)
    singles = []
    for (m, n, c, r, vv) in (mats if sequential else []):
        P = fl.FastLoadMatrix(m, n, c, r, vv, width=32)
        P.prepare(fl.VAL_F32)
        Q = fl.FastLoadMatrix(m, n, c, r, vv, width=64)
        Q.prepare_rows()
        S = Q if sched.width == 64 else fl.FastLoadMatrix(m, n, c, r, vv, width=sched.width)
        prep_sched(S)
        singles.append((P, Q, S, P.perm(), S.perm() if sched.kernel >= 0 else None))
    log(f"[batched] per-matrix preprocessing done ({time.time() - t0:.1f}s)")
    for k in ks:
        X = rng.standard_normal((N_, k)).astype(np.float32)
        dX = DeviceArray.from_host(X.ravel())
        dY = DeviceArray(M_ * k, np.float32)
        B_alg = nnz * 8 + N_ * k * 4 + M_ * k * 4
        if k == 1 and sched.kernel >= 0:
            dsx = DeviceArray(N_, np.float32)
            B1.permute_x(dX, dsx)

            def batch_step():
                B1.spmv(dsx, dY, sched.kernel, fl.VAL_F32, stream=s, assign=True)
        else:
            def batch_step():
                B.spmm(dX, dY, k, fl.VAL_F32, stream=s)
        batch_step()
        s.synchronize()
        y_batch = dY.to_host().reshape(M_, k)
        g = Graph(s)
        with g:
            batch_step()
        t_graph = timed(g.launch, steps)
        t_eager = timed(batch_step, steps)
        if not sequential:
            res[f"k{k}"] = dict(graph_ms=round(t_graph, 4), eager_ms=round(t_eager, 4),
                                graph_gbs=round(B_alg / t_graph / 1e6, 1), bytes_alg=B_alg)
            log(f"[batched] k={k:2d} graph {t_graph:.3f} ms  eager {t_eager:.3f} ms  ({res[f'k{k}']['graph_gbs']} GB/s)")
            del g
            continue
        # plain drop-in, one launch per matrix and right-hand side: x columns permuted on the host,
        # matrix i's column j at offset co[i] * k + j * n_i (likewise y)
        XP = np.empty(N_ * k, np.float32)
        for i, (P, Q, S, perm, _) in enumerate(singles):
            XP[co[i] * k:co[i + 1] * k] = X[co[i]:co[i + 1]][perm].T.ravel()
        dXP = DeviceArray.from_host(XP)
        dYP = DeviceArray(M_ * k, np.float32)
        plain = []
        for i, (P, Q, S, perm, _) in enumerate(singles):
            mi, ni = mats[i][0], mats[i][1]
            for j in range(k):
                plain.append((P, dXP.view(co[i] * k + j * ni, ni), dYP.view(ro[i] * k + j * mi, mi)))

        def seq_plain():
            for P, sx, y in plain:
                fl.zero_device(y, s)
                P.spmv(sx, y, fl.KERNEL_DROPIN, fl.VAL_F32, stream=s)
        t_plain = timed(seq_plain, max(steps // 10, 3), warm=min(warmup, 2))
        YP = dYP.to_host()
        y_plain = np.concatenate([YP[ro[i] * k:ro[i + 1] * k].reshape(k, -1).T for i in range(count)])
        # the batch's own schedule per matrix
        same = []
        for i, (P, Q, S, perm, sperm) in enumerate(singles):
            mi, ni = mats[i][0], mats[i][1]
            xi = dX.view(co[i] * k, ni * k)
            yi = dY.view(ro[i] * k, mi * k)
            if k == 1 and sched.kernel >= 0:
                sxi = DeviceArray.from_host(X[co[i]:co[i + 1], 0][sperm])
                same.append((S, sxi, yi, True))
            else:
                same.append((Q, xi, yi, False))

        def seq_same():
            for S, x_, y_, spmv in same:
                if spmv:
                    S.spmv(x_, y_, sched.kernel, fl.VAL_F32, stream=s, assign=True)
                else:
                    S.spmm(x_, y_, k, fl.VAL_F32, stream=s)
        t_same = timed(seq_same, max(steps // 10, 3), warm=min(warmup, 2))
        y_same = dY.to_host().reshape(M_, k)
        nb = max(np.linalg.norm(y_plain), 1e-300)
        err_plain = float(np.linalg.norm(y_batch.astype(np.float64) - y_plain) / nb)
        same_bitwise = bool(np.array_equal(y_batch.view(np.uint32), y_same.view(np.uint32)))
        err_same = float(np.linalg.norm(y_batch.astype(np.float64) - y_same) / nb)
        sched_name = (f"{sched.name} (learned)" if k == 1 else "row")
        res[f"k{k}"] = dict(graph_ms=round(t_graph, 4), eager_ms=round(t_eager, 4), graph_gbs=round(B_alg / t_graph / 1e6, 1),
                            bytes_alg=B_alg, schedule=f"{sched_name} on the block-diagonal batch",
                            sequential_plain_ms=round(t_plain, 4),
                            sequential_plain=f"plain FP32 drop-in, {count * k} launches ({count} matrices x {k} "
                                             f"right-hand sides), eager",
                            speedup_vs_sequential_plain=round(t_plain / t_graph, 2),
                            rel_diff_vs_sequential_plain=err_plain,
                            sequential_same_ms=round(t_same, 4),
                            sequential_same=f"{sched_name} per matrix, {count} launches, eager",
                            speedup_vs_sequential_same=round(t_same / t_graph, 2),
                            same_bit_identical=same_bitwise, rel_diff_vs_sequential_same=err_same)
        log(f"[batched] k={k:2d} graph {t_graph:.3f} ms  eager {t_eager:.3f} ms  ({res[f'k{k}']['graph_gbs']} GB/s)  "
            f"plain x{k} seq {t_plain:.3f} ms (x{res[f'k{k}']['speedup_vs_sequential_plain']}, diff {err_plain:.1e})  "
            f"same-schedule seq {t_same:.3f} ms (x{res[f'k{k}']['speedup_vs_sequential_same']}, "
            f"{'bit-identical' if same_bitwise else f'diff {err_same:.1e}'})")
        del g, plain, same
    return res


def run_cg(fl, steps, warmup, g=128, iters=200):
    """Conjugate gradients (fl_cg_*) on the resident matrix, FP64: the symmetric positive definite
    27-point matrix on g^3 (26 on the diagonal, -1 elsewhere; the stencil workload's structure)
    with the stencil's schedule (WINR, 64-tile batches). One solve to ||r||/||b|| <= 1e-8 from
    x = 0 (iterations, time, residual recomputed on the host), then `iters` iterations with
    rtol = 0, timed on the host around the solve after a warm-up solve (the graph of 50
    iterations is captured by the first solve and replayed; the host reads the residual between
    replays). Bytes per iteration: the SpMV's algorithmic bytes in FP64 plus 13 vector passes of
    n doubles (p permuted to sorted order: 2; p.q: 2; x += alpha p, r -= alpha q: 6; p = r + beta
    p: 3). The SpMV alone is timed with the same schedule for comparison."""
    import scipy.sparse as sp
    from fastload_amd import DeviceArray
    from fastload_amd import workloads as wl
    m, n, cp, ri, v, b = wl.laplacian27(g)
    nnz = int(cp[n])
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare_schedule(fl.KERNEL_WINR, 64, 0)
    cg = fl.CG(M, fl.KERNEL_WINR, fl.VAL_F64)
    db = DeviceArray.from_host(b)
    dx = DeviceArray.zeros(n, np.float64)
    every = 50
    t0 = time.perf_counter()
    it_conv, res_conv = cg.solve(db, dx, max_iter=5000, rtol=1e-8, check_every=every)
    conv_s = time.perf_counter() - t0
    A = sp.csc_matrix((v, ri, cp), shape=(m, n))
    x = dx.to_host()
    true_res = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    del A, x
    for _ in range(max(warmup, 1)):
        dx = DeviceArray.zeros(n, np.float64)
        cg.solve(db, dx, max_iter=iters, rtol=0.0, check_every=every)
    fl.synchronize()
    dx = DeviceArray.zeros(n, np.float64)
    t0 = time.perf_counter()
    it, _ = cg.solve(db, dx, max_iter=iters, rtol=0.0, check_every=every)
    fl.synchronize()
    ms_iter = (time.perf_counter() - t0) * 1e3 / max(it, 1)
    # the SpMV alone: y = A sortx with the same schedule
    dsx, dy = DeviceArray.zeros(n, np.float64), DeviceArray.zeros(m, np.float64)
    M.permute_x(DeviceArray.from_host(b), dsx)
    spmv_step, spmv_kernel, _ = time_variant(
        fl, lambda: M.spmv(dsx, dy, fl.KERNEL_WINR, fl.VAL_F64, assign=True), dy, steps, warmup, zero=False)
    spmv_bytes = b_alg(nnz, n, m, 8, 8, 8)
    it_bytes = spmv_bytes + 13 * n * 8
    cg.close()
    r = {"workload": f"27-point SPD matrix on {g}^3 (26 on the diagonal, -1 elsewhere), b N(0,1), FP64",
         "n": n, "nnz": nnz, "schedule": "winr64 (WINR, 64-tile batches)", "graph_iterations": every,
         "converged_iterations": it_conv, "converged_rel_res": res_conv, "true_rel_res": true_res,
         "converge_s": round(conv_s, 4), "timed_iterations": it, "ms_per_iter": round(ms_iter, 5),
         "bytes_per_iter": it_bytes, "gbs": round(it_bytes / ms_iter / 1e6, 1),
         "frac_of_peak": round(it_bytes / ms_iter / 1e6 / PEAK_GBS, 4),
         "spmv_step_ms": round(spmv_step, 5), "spmv_kernel_ms": round(spmv_kernel, 5),
         "spmv_share": round(spmv_step / ms_iter, 3)}
    log(f"[cg] {g}^3 FP64: {it_conv} iterations to 1e-8 in {conv_s:.3f} s (true res {true_res:.2e}); "
        f"{ms_iter:.4f} ms/iteration = {r['gbs']} GB/s; SpMV alone {spmv_step:.4f} ms")
    return r


def run_cache(fl, g=128):
    """The preprocessing cache file (fl_matrix_save / fl_matrix_load) against the preparation it
    replaces, on the 27-point SPD matrix on g^3 with the WINR schedule (64-tile batches): wall
    time of building the device layout from host CSC arrays plus the schedule, of saving, and of
    loading (twice: the second read finds the file in the page cache), and how far the loaded
    matrix's y is from the original's, beside the spread of two runs of the original (WINR in
    FP64 flushes rows that several batches touch with atomics, so y is reproducible only up to
    the order of those additions)."""
    import tempfile
    from fastload_amd import DeviceArray
    from fastload_amd import workloads as wl
    m, n, cp, ri, v, b = wl.laplacian27(g)
    fl.synchronize()
    t0 = time.perf_counter()
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare_schedule(fl.KERNEL_WINR, 64, 0)
    fl.synchronize()
    prep_s = time.perf_counter() - t0
    fd, path = tempfile.mkstemp(suffix=".flm")
    os.close(fd)
    try:
        t0 = time.perf_counter()
        M.save(path)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(path)
        loads = []
        for _ in range(2):
            t0 = time.perf_counter()
            L = fl.FastLoadMatrix.load(path)
            fl.synchronize()
            loads.append(time.perf_counter() - t0)
    finally:
        os.remove(path)
    dsx = DeviceArray.zeros(n, np.float64)
    M.permute_x(DeviceArray.from_host(b), dsx)
    ys = []
    for X in (M, M, L):
        dy = DeviceArray.zeros(m, np.float64)
        X.spmv(dsx, dy, fl.KERNEL_WINR, fl.VAL_F64, assign=True)
        ys.append(dy.to_host())
    scale = max(float(np.abs(ys[0]).max()), np.finfo(np.float64).tiny)
    r = {"workload": f"27-point SPD matrix on {g}^3, nnz {int(cp[n])}, layout W=64 + WINR schedule (64-tile batches)",
         "prepare_s": round(prep_s, 4), "save_s": round(save_s, 4), "file_bytes": size,
         "load_s_first": round(loads[0], 4), "load_s_cached": round(loads[1], 4),
         "loaded_vs_original_max_diff": float(np.abs(ys[2] - ys[0]).max() / scale),
         "original_run_to_run_max_diff": float(np.abs(ys[1] - ys[0]).max() / scale)}
    log(f"[cache] prepare {prep_s:.3f} s, save {save_s:.3f} s ({size / 1e9:.2f} GB), load {loads[0]:.3f} s "
        f"/ {loads[1]:.3f} s, y diff {r['loaded_vs_original_max_diff']:.1e} "
        f"(run to run {r['original_run_to_run_max_diff']:.1e})")
    return r


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_threads() -> int:
    """Cores this process may use, capped at the 16 a one-GPU host grants a job."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(M, m, n, nnz, budget_s=10.0):
    """The CPU path over the same 32-wide layout, y zeroed before each pass as a GPU step does:
      - the reference's own FastLoad_cpu (src/FastLoad_CPU.h:2-46), compiled from the reference
        sources into oracle/_ref, FP64 (its only precision), one core, about budget_s seconds
        (falls back to the oracle's serial FP32 port when oracle/_ref was not built);
      - this library's fl_FastLoad_cpu on every core the process may use (tiles split into
        ranges, atomic y), FP64, about budget_s seconds.
    `value` is in the headline's unit, GB/s-equivalent with FP32 bytes (nnz*8 + n*4 + m*4)."""
    import ctypes as C
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import oracle_bind as ob
    e = M.export()
    x = np.random.default_rng(0).standard_normal(n)
    B32 = b_alg(nnz, n, m, 4, 4, 4)
    sx = np.ascontiguousarray(x[e["perm"]], np.float64)
    L = ob.ref_lib(32) if ob.ref_available(32) else None
    reps, t_total = 0, 0.0
    out = None
    if L is not None:
        keep = dict(sortval=np.ascontiguousarray(e["sortval"], np.float64),
                    sortrowidx=np.ascontiguousarray(e["sortrowidx"], np.int32),
                    tile_ptr=np.ascontiguousarray(e["tile_ptr"], np.int32),
                    tile_len=np.ascontiguousarray(e["tile_len"], np.int32),
                    tile_colidx=np.ascontiguousarray(e["tile_colidx"], np.int32))
        sm = ob.SlideMatrix()
        sm.sortval = keep["sortval"].ctypes.data_as(ob._DP)
        for k in ("sortrowidx", "tile_ptr", "tile_len", "tile_colidx"):
            setattr(sm, k, keep[k].ctypes.data_as(ob._IP))
        sm.tilenum = int(e["tilenum"])
        y = np.zeros(m)
        while t_total < budget_s and reps < 100:
            t0 = time.perf_counter()
            y[:] = 0.0
            L.ref_fastload_cpu(C.byref(sm), nnz, m, n, sx, y)
            t_total += time.perf_counter() - t0
            reps += 1
        per = t_total / reps
        y_ref1 = y.copy()
        out = dict(value=round(B32 / per / 1e9, 3), unit="GB/s", cores=1, kind="reference",
                   sample=f"{reps} full FastLoad_cpu passes (reference source, FP64, y zeroed each pass) over the "
                          f"same 32-wide layout, {per * 1e3:.1f} ms each; GB/s-equivalent with FP32 bytes",
                   ms_per_spmv=round(per * 1e3, 3), fp64_bytes_gbs=round(b_alg(nnz, n, m, 8, 8, 8) / per / 1e9, 3))
    else:
        t = dict(tilenum=e["tilenum"], tile_ptr=e["tile_ptr"], tile_len=e["tile_len"], tile_colidx=e["tile_colidx"],
                 sortrowidx=e["sortrowidx"])
        v32 = e["sortval"].astype(np.float32)
        x32 = sx.astype(np.float32)
        while t_total < budget_s and reps < 100:
            t0 = time.perf_counter()
            ob.fastload_cpu_f32(t, v32, x32, m)
            t_total += time.perf_counter() - t0
            reps += 1
        per = t_total / reps
        y_ref1 = None
        out = dict(value=round(B32 / per / 1e9, 3), unit="GB/s", cores=1, kind="port",
                   sample=f"{reps} full FastLoad_cpu FP32 passes (oracle port) over the same 32-wide layout "
                          f"({per * 1e3:.1f} ms each)", ms_per_spmv=round(per * 1e3, 3))
    out["cpu"] = cpu_model()
    # this library's host path on every available core
    import fastload_amd as fl
    tiles = fl.Tiles(int(e["tilenum"]), int(e["segsum"]), e["tile_ptr"], e["tile_len"], e["tile_colidx"],
                     e["tile_format"], e["sortrowidx"], e["sortval"], e["sortrowindex"], e["countrow"],
                     e["segmentoffset"])
    T = cpu_threads()
    y = np.zeros(m)
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 100:
        t0 = time.perf_counter()
        y[:] = 0.0
        fl.FastLoad_cpu(tiles, nnz, m, n, sx, threads=T, y=y)
        t_total += time.perf_counter() - t0
        reps += 1
    per = t_total / reps
    diff = float(np.linalg.norm(y - y_ref1) / max(np.linalg.norm(y_ref1), 1e-300)) if y_ref1 is not None else None
    out["all_cores"] = dict(value=round(B32 / per / 1e9, 3), unit="GB/s", cores=T, ms_per_spmv=round(per * 1e3, 3),
                            impl="fl_FastLoad_cpu (this library): tile ranges per thread, atomic y updates, FP64",
                            sample=f"{reps} full passes, y zeroed each pass", rel_diff_vs_1core=diff)
    log(f"[cpu] 1 core {out['ms_per_spmv']} ms/SpMV ({out['value']} GB/s); {T} cores {per * 1e3:.1f} ms/SpMV "
        f"({out['all_cores']['value']} GB/s) on {out['cpu']}")
    return out


PUBLISHED_1138_MS = 0.066730  # src/TimeResult/results_slidecuda.txt:1 (RTX 3090 Ti per readme.md:18)


def run_c1(fl, steps, warmup):
    """Plain FastLoad FP64 on 1138_bus (values i % 10 and x i % 10, as src/main.cu:68-69,93-96 set
    them) and on the seeded structurally symmetric twin. Each goes through the reference's host API
    (mmio_allinone / csr_to_csc / ColSort / formatTransform / FastLoad_cpu / FastLoad_spmv): the
    reference measures one FastLoad_spmv launch as the host clock around launch + synchronize,
    mean of 100 (src/FastLoad_GPU.h:349-374), and that is the number it publishes for 1138_bus
    (0.066730 ms). The device-resident drop-in kernel is also timed with HIP events."""
    from fastload_amd import DeviceArray
    from fastload_amd import workloads as wl
    root = os.path.dirname(os.path.abspath(__file__))
    cases = {}
    m, n, nnz, _, rp, ci, v = fl.mmio_allinone(os.path.join(root, "tests", "golden", "1138_bus.mtx"))
    v = (np.arange(nnz) % 10).astype(np.float64)
    cases["1138_bus"] = (m, n, rp, ci, v, (np.arange(n) % 10).astype(np.float64))
    m2, n2, cp2, ri2, v2, x2 = wl.symgeom()
    import scipy.sparse as sp
    A = sp.csc_matrix((v2, ri2, cp2), shape=(m2, n2)).tocsr()
    A.sort_indices()
    cases["symgeom_twin"] = (m2, n2, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, x2)
    out = {}
    for name, (m, n, rp, ci, v, x) in cases.items():
        cv, cr, cp, npc, _ = fl.csr_to_csc(m, n, v, rp, ci)
        tr, tv, sn, sx, _ = fl.ColSort(m, n, npc, cp, cr, cv, x)
        tiles, h, _, _ = fl.formatTransform(tr, tv, sn, len(cv), m, n)
        y_gold = np.zeros(m)
        for c in range(n):  # y_goldencsc, src/main.cu:110-120
            y_gold[cr[cp[c]:cp[c + 1]]] += cv[cp[c]:cp[c + 1]] * x[c]
        t0 = time.perf_counter()
        y_cpu, cpu_err = fl.FastLoad_cpu(tiles, len(cv), m, n, sx, y_golden=y_gold if name == "1138_bus" else None)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        y, wall_ms, passed = fl.FastLoad_spmv(tiles, len(cv), n, m, sx, y_golden=y_gold)
        M = fl.FastLoadMatrix(m, n, cp, cr, cv, width=32)
        dsx = DeviceArray.from_host(sx)
        dy = DeviceArray(m, np.float64)
        _, kern_ms, kern_med = time_variant(fl, lambda: M.spmv(dsx, dy, fl.KERNEL_DROPIN, fl.VAL_F64), dy,
                                            steps, warmup)
        rel = float(np.linalg.norm(dy.to_host() - y_gold) / max(np.linalg.norm(y_gold), 1e-300))
        out[name] = dict(m=m, n=n, nnz=int(len(cv)), levels=h, tiles=tiles.tilenum, fp64=True,
                         fastload_spmv_ms_reference_method=round(wall_ms, 5), check="PASS" if passed else "NO PASS",
                         kernel_event_ms_mean=round(kern_ms, 5), kernel_event_ms_median=round(kern_med, 5),
                         rel_err_vs_golden=rel, fastload_cpu_ms=round(cpu_ms, 4),
                         fastload_cpu_exact=(cpu_err == 0) if name == "1138_bus" else None)
        if name == "1138_bus":
            out[name].update(published_ms=PUBLISHED_1138_MS, published_hw="RTX 3090 Ti (readme.md:18), wall clock",
                             speedup_vs_published=round(PUBLISHED_1138_MS / wall_ms, 2))
        log(f"[c1] {name}: FastLoad_spmv {wall_ms:.4f} ms (reference method), kernel {kern_ms * 1e3:.1f} us "
            f"(events), FastLoad_cpu {cpu_ms:.3f} ms, {out[name]['check']}")
    return out


def run_all(args):
    """Every workload bench.py times; returns the JSON object."""
    import fastload_amd as fl
    from fastload_amd import DeviceArray
    from fastload_amd import workloads as wl

    flush = None if args.no_flush else DeviceArray(1 << 28, np.float32)  # 1 GiB, 4x the Infinity Cache
    out, comparators = {"profiling_run": "the power-law workload was skipped (--no-powerlaw)"}, {}
    if not args.no_powerlaw:
        t0 = time.time()
        m, n, cp, ri, v, x = wl.powerlaw(args.n, args.seed)
        log(f"[powerlaw] generated {m}x{n} nnz={int(cp[n])} in {time.time() - t0:.1f}s")
        dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
        del cp, ri, v
        sched, chosen = select_schedule(fl, m, n, dcp, dri, dv, force=args.schedule)
        log(f"[powerlaw] selector chose {chosen}")
        variants = {k: VARIANTS[k] for k in args.variants.split(",") if k}
        variants.update(learned_variants(sched, args.learned_kinds.split(",")))
        nnz, mats, res, y_ref = run_workload(fl, "powerlaw", m, n, dcp, dri, dv, x, variants, args.steps, args.warmup,
                                             flush)
        del mats
        comparators = {} if args.no_compare else {
            "powerlaw": run_comparators(fl, "powerlaw", m, n, dcp, dri, dv, x, y_ref, args.steps, args.warmup, flush)}
        del dcp, dri, dv

        ok = {k: r for k, r in res.items() if r["rel_err_l2"] <= 1e-4 and k.startswith("learned")}
        best = max(ok, key=lambda k: ok[k]["gbs_equiv_f32"])
        r = res[best]
        plain = res.get("plain_f32")
        kind_names = {"f32": "f32", "f64": "f64", "bf16": "bf16 values, f32 products, f64 accumulate",
                      "bf16x2": "split hi+lo bf16 values, f32 products, f64 accumulate",
                      "packed": "packed value+slot words (value rounded to the leftover mantissa bits), f32 products, "
                                "f64 accumulate"}
        out = {
            "metric": "SpMV HBM GB/s (% of peak): learned+mixed-prec vs plain FastLoad at stated err",
            "value": r["gbs_equiv_f32"],
            "unit": "GB/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": r["step_ms"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": kind_names[best.split("_", 1)[1]],
            "data": "synthetic",
            "config": {"workload": f"power-law CSC SpMV {m}x{n}, Zipf({wl.POWERLAW_ZIPF}, max {wl.POWERLAW_MAX}) column "
                                   f"nnz, seed {args.seed}",
                       "nnz": nnz, "variant": best, "learned_schedule": chosen, "rel_err_l2_vs_fp64": r["rel_err_l2"],
                       "plain_f32_gbs": plain["gbs"] if plain else None,
                       "speedup_vs_plain_f32": round(plain["step_ms"] / r["step_ms"], 3) if plain else None},
            "roofline": {"bound": "hbm", "achieved": r["kernel_gbs_equiv_f32"], "peak": PEAK_GBS, "unit": "GB/s",
                         "frac": round(r["kernel_gbs_equiv_f32"] / PEAK_GBS, 4),
                         "traffic": profiled_traffic(chosen["schedule"], best.split("_", 1)[1], "powerlaw"),
                         "flushed_achieved": r.get("kernel_gbs_equiv_f32_flushed"),
                         "flushed_frac": (round(r["kernel_gbs_equiv_f32_flushed"] / PEAK_GBS, 4)
                                          if "kernel_gbs_equiv_f32_flushed" in r else None)},
            "variants": res,
        }
        if comparators:
            lf = r["kernel_ms"]
            for k, c in comparators["powerlaw"].items():
                c["learned_speedup"] = round(c["kernel_ms"] / lf, 2)
    if not args.no_stencil:
        sm, sn, scp, sri, sv, sx = wl.stencil(128)
        dcp, dri, dv = DeviceArray.from_host(scp), DeviceArray.from_host(sri), DeviceArray.from_host(sv)
        del scp, sri, sv
        ssched, schosen = select_schedule(fl, sm, sn, dcp, dri, dv, force=args.stencil_schedule)
        log(f"[stencil] selector chose {schosen}")
        svars = {k: VARIANTS[k] for k in args.stencil_variants.split(",") if k}
        svars.update(learned_variants(ssched, args.learned_kinds.split(",")))
        svars.update({k: VARIANTS[k] for k in ("win_bf16x2", "win_bf16")})
        snnz, smats, sres, sy_ref = run_workload(fl, "stencil", sm, sn, dcp, dri, dv, sx, svars, args.steps,
                                                 args.warmup, flush)
        del smats
        if not args.no_compare:
            comparators["stencil"] = run_comparators(fl, "stencil", sm, sn, dcp, dri, dv, sx, sy_ref, args.steps,
                                                     args.warmup, flush)
            slf = sres["learned_f32"]["kernel_ms"] if "learned_f32" in sres else None
            for c in comparators["stencil"].values():
                c["learned_speedup"] = round(c["kernel_ms"] / slf, 2) if slf else None
        out["stencil"] = {"workload": "27-point stencil on 128^3, |v| U[0.5,1.5] random sign, seed 128", "nnz": snnz,
                          "learned_schedule": schosen,
                          "traffic_bytes_learned_f32": profiled_traffic(schosen["schedule"], "f32", "stencil"),
                          "variants": sres}
        del dcp, dri, dv
    del flush
    if comparators:
        out["comparators"] = comparators
    if not args.no_c1:
        out["c1"] = run_c1(fl, args.steps, args.warmup)
    if not args.no_cg:
        out["cg"] = run_cg(fl, args.steps, args.warmup)
        try:  # writes a 1.3 GB temporary file: a full disk costs this entry, not the bench line
            out["cache_file"] = run_cache(fl)
        except (OSError, fl.FastLoadError) as e:
            log(f"[cache] skipped: {e}")
            out["cache_file"] = {"error": str(e)}
    if args.batch > 0:
        out["batched"] = run_batched(fl, args.batch, 512, args.steps, args.warmup, sequential=not args.no_seq,
                                     ks=tuple(int(k) for k in args.batch_k.split(",")))
    if not args.no_cpu and not args.no_powerlaw:
        m, n, cp, ri, v, _ = wl.powerlaw(args.n, args.seed)
        M32 = fl.FastLoadMatrix(m, n, cp, ri, v, width=32)
        del cp, ri, v
        out["cpu_baseline"] = cpu_baseline(M32, m, n, nnz)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stencil", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-cg", action="store_true", help="skip the conjugate-gradient solve and the cache file")
    ap.add_argument("--no-flush", action="store_true", help="skip the Infinity-Cache-flushed timings")
    ap.add_argument("--no-compare", action="store_true", help="skip the naive-CSC and rocSPARSE yardsticks")
    ap.add_argument("--batch", type=int, default=512, help="matrices in the batched configuration (0 = skip)")
    ap.add_argument("--batch-k", default="1,8,32", help="right-hand-side counts of the batched configuration")
    ap.add_argument("--no-seq", action="store_true", help="skip the batched configuration's sequential baselines")
    ap.add_argument("--no-powerlaw", action="store_true", help="skip the power-law workload (profiling runs)")
    ap.add_argument("--variants", default="plain_f32,plain_f64,fast_f32,win_f32,pb_f32,pb_f64,row_f32,hyb_f32,hyb_f64")
    ap.add_argument("--stencil-variants", default="plain_f32,fast_f32,win_f32,win_f64,winr64_f32,pb_f32,row_f32")
    ap.add_argument("--schedule", default=None, help="use this schedule for the learned variants instead of "
                    "selecting one (profiling runs)")
    ap.add_argument("--stencil-schedule", default=None, help="the same for the stencil workload")
    ap.add_argument("--learned-kinds", default="f32,bf16x2,bf16,packed",
                    help="value storages timed for the learned schedule")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this FastLoad build runs on one GPU only (no multi-GPU path is built); "
                                   "run with --gpus 1"}))
        sys.exit(2)

    # The library's reference-style host API prints its check lines ("Check FASTLOAD GPU PASS!") on
    # stdout; the one JSON line must be the only thing there, so fd 1 points at stderr while the
    # workloads run.
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        out = run_all(args)
    finally:
        sys.stdout.flush()
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
