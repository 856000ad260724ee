"""Learned schedule selector: 40 matrix features computed on the GPU -> a 40-64-64-13 perceptron
trained and run by the library's own HIP kernels (fl_mlp_*: MFMA with BF16 weights and FP32
accumulation) -> one of 13 SpMV schedules.

The reference has one fixed schedule (the 32-wide tiles of src/formatTransform_GPU.h with the
segmented-sum kernel of src/FastLoad_GPU.h); that schedule is class 0, the selector's default and
fallback, and the yardstick of its regret. Its column statistics (src/auxiliary.h:6-25: mean,
standard deviation, ratio of column lengths) are computed and never used there; here a
histogram of column lengths plus row statistics drive the choice.

Schedule space (tile width x reduction strategy x batch/band parameters):
  0 dropin_w32   reference tiles and segmented sums (32 lanes per tile), atomics per segment
  1 dropin_w64   the same algorithm on 64-wide tiles
  2 winr64       WIN on 64-tile batches of tiles grouped by row window (owned rows, no atomic)
  3-5 win{64,128,256}  batches of 64/128/256 tiles, rows compacted to LDS slots, one flush per
                 row of a batch (a store for rows no other batch touches, else one atomic)
  6-10 pb*       propagation blocking, batch 128/256 tiles x band 4K/8K/16K rows, no atomics
  11 row         entries re-ordered by row (CSR), four lanes per row, no atomics (fl_spmm, k=1)
  12 hyb32k_pb128_r16k  the min(n, 32768) longest columns row by row with their x in LDS, the
                 rest PB 128 tiles x 16K rows on a sub-matrix

Training data come from exhaustive sweeps: every schedule is timed on every matrix of a seeded
synthetic population (`population_spec`), with the matrix and x resident on the GPU; one step is
y = A x in FP32 (`fl_spmv_assign`, or `fl_spmm` for `row`; value storage FP32, FP64 LDS sums
where the schedule has them). Preprocessing time is not part of a step (it is paid once per
matrix). On a new matrix the network's probabilities pick the candidates that are timed before
the fastest is kept (`Selector.choose_refined`, `tuning_candidates`); the HYB schedule is timed
as well whenever the 32768 longest columns hold at least 35 % of the entries.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._lib import check, lib
from .api import (GEN_BANDED, GEN_BLOCKDIAG, GEN_STENCIL27, GEN_UNIFORM, GEN_ZIPF, GEN_ZIPF_DENSE, KERNEL_DROPIN,
                  KERNEL_FAST,
                  KERNEL_HYB,
                  KERNEL_PB, KERNEL_WIN, KERNEL_WINR, VAL_F32, FastLoadMatrix, gen_matrix)
from .device import DeviceArray, Event, synchronize

N_FEATURES = 40
HIDDEN = (64, 64)
KERNEL_ROW = -1  # fl_spmm with k = 1


@dataclass(frozen=True)
class Schedule:
    name: str
    width: int
    kernel: int
    batch_tiles: int = 0
    band_rows: int = 0


HYB_COLS = 32768  # heavy columns of the HYB class: min(n, HYB_COLS)


def _schedule_table():
    """The 13 schedules, read from the library (fl_schedule_info) so C and Python callers share
    one table; names follow the parameters (dropin_w32, win128, pb256_r16k, row, ...)."""
    kname = {KERNEL_DROPIN: "dropin", KERNEL_FAST: "fast", KERNEL_WIN: "win", KERNEL_WINR: "winr", KERNEL_PB: "pb",
             KERNEL_ROW: "row", KERNEL_HYB: "hyb"}
    out = []
    for c in range(FL_SCHEDULES):
        w, k, bt, r = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib.fl_schedule_info(c, C.byref(w), C.byref(k), C.byref(bt), C.byref(r)), "fl_schedule_info")
        w, k, bt, r = w.value, k.value, bt.value, r.value
        if k in (KERNEL_DROPIN, KERNEL_FAST):
            name = f"{kname[k]}_w{w}"
        elif k in (KERNEL_WIN, KERNEL_WINR):
            name = f"{kname[k]}{bt}"
        elif k == KERNEL_PB:
            name = f"pb{bt}_r{r // 1024}k"
        elif k == KERNEL_HYB:
            name = f"hyb{HYB_COLS // 1024}k_pb{bt}_r{r // 1024}k"
        else:
            name = "row"
        out.append(Schedule(name, w, k, bt, r))
    return tuple(out)


FL_SCHEDULES = 13
SCHEDULES = _schedule_table()
N_CLASSES = len(SCHEDULES)
DEFAULT_CLASS = 0
HYB_CLASS = next(c for c, sc in enumerate(SCHEDULES) if sc.kernel == KERNEL_HYB)
# choose_refined also times the HYB class when the HYB_COLS longest columns hold at least
# HYB_MIN_SHARE of the entries, whatever probability the network gives it
HYB_MIN_SHARE = 0.35


def heavy_share(n, dcolptr, cols: int = HYB_COLS) -> float:
    """Share of the entries held by the `cols` longest columns of a device CSC matrix."""
    cp = dcolptr.to_host()[:n + 1].astype(np.int64)
    nnz = int(cp[n]) if n else 0
    if nnz == 0:
        return 0.0
    cnt = np.diff(cp)
    k = min(cols, n)
    return float(np.partition(cnt, n - k)[n - k:].sum()) / nnz

DATA_DIR = Path(__file__).resolve().parent / "data"
DATASET_PATH = DATA_DIR / "selector_sweep.npz"
# Model-guided tuning times every schedule the network gives at least TUNE_MIN_PROB probability
# (the TUNE_MAX most likely at most, the top one always). On the training matrices 2 % gives a
# lower regret than timing the top three (1.0002 against 1.0006) with fewer timings (2.1).
TUNE_MIN_PROB = 0.02
TUNE_MAX = 5
WEIGHTS_PATH = DATA_DIR / "selector_weights.npz"


# ------------------------------------------------------------------------------ features
def features(m: int, n: int, dcolptr: DeviceArray, drowidx: DeviceArray, stream=None) -> np.ndarray:
    """The 40 selector features of a device CSC matrix (include/fastload.h fl_features)."""
    out = np.zeros(N_FEATURES, np.float32)
    nnz = drowidx.count if drowidx is not None else 0
    check(lib.fl_features(m, n, nnz, dcolptr.ptr, drowidx.ptr if nnz else None, out.ctypes.data,
                          stream.ptr if stream is not None else None), "fl_features")
    return out


# ------------------------------------------------------------------------------------ MLP
class MLP:
    """The 40-64-64-13 perceptron of fl_mlp_* (forward, backward and Adam on the GPU)."""

    def __init__(self, dims=(N_FEATURES, *HIDDEN, N_CLASSES), seed: int = 0):
        h = C.c_void_p()
        check(lib.fl_mlp_create(C.byref(h), *dims, seed), "fl_mlp_create")
        self.handle = h.value
        self.dims = tuple(dims)
        self.count = lib.fl_mlp_param_count(self.handle)

    def shapes(self):
        d = self.dims
        return [(d[0], d[1]), (d[1],), (d[1], d[2]), (d[2],), (d[2], d[3]), (d[3],)]

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.count, np.float32)
        check(lib.fl_mlp_get_params(self.handle, p.ctypes.data), "fl_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.count:
            raise ValueError(f"expected {self.count} parameters, got {p.size}")
        check(lib.fl_mlp_set_params(self.handle, p.ctypes.data), "fl_mlp_set_params")

    def unpack(self, flat):
        out, o = [], 0
        for s in self.shapes():
            k = int(np.prod(s))
            out.append(np.asarray(flat[o:o + k]).reshape(s))
            o += k
        return out

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        dX = DeviceArray.from_host(X.ravel())
        dL = DeviceArray(X.shape[0] * self.dims[-1], np.float32)
        check(lib.fl_mlp_forward(self.handle, dX.ptr, X.shape[0], dL.ptr, None), "fl_mlp_forward")
        return dL.to_host().reshape(X.shape[0], self.dims[-1])

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.dims[0])
        dX = DeviceArray.from_host(X.ravel())
        cls = np.zeros(X.shape[0], np.int32)
        prob = np.zeros(X.shape[0], np.float32)
        check(lib.fl_mlp_predict(self.handle, dX.ptr, X.shape[0], cls.ctypes.data, prob.ctypes.data, None),
              "fl_mlp_predict")
        return cls, prob

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        dX, dT = DeviceArray.from_host(X.ravel()), DeviceArray.from_host(T.ravel())
        g = np.zeros(self.count, np.float32)
        loss = C.c_float()
        check(lib.fl_mlp_gradients(self.handle, dX.ptr, dT.ptr, X.shape[0], g.ctypes.data, C.byref(loss), None),
              "fl_mlp_gradients")
        return g, float(loss.value)

    def train_step(self, x_ptr: int, t_ptr: int, batch: int, lr: float) -> float:
        """One Adam step on `batch` rows at device addresses x_ptr (features) and t_ptr (targets);
        returns the loss before the step."""
        loss = C.c_float()
        check(lib.fl_mlp_train_step(self.handle, x_ptr, t_ptr, batch, lr, C.byref(loss), None), "fl_mlp_train_step")
        return float(loss.value)

    def close(self):
        if getattr(self, "handle", None):
            lib.fl_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------- targets
def soft_targets(times: np.ndarray, tau: float = 0.05) -> np.ndarray:
    """Distribution over schedules from measured step times: p_c ~ exp(-ln(t_c / t_best) / tau),
    so a schedule 5% slower than the best gets weight e^-1 relative to it. Missing times (inf/nan)
    get zero weight."""
    t = np.asarray(times, np.float64)
    ok = np.isfinite(t) & (t > 0)
    lt = np.where(ok, np.log(np.where(ok, t, 1.0)), np.inf)
    best = lt.min(axis=1, keepdims=True)
    w = np.where(ok, np.exp(-(lt - best) / tau), 0.0)
    return (w / w.sum(axis=1, keepdims=True)).astype(np.float32)


def train(F, times, idx, steps=3000, batch=128, lr=3e-3, tau=0.05, seed=0, log_every=0):
    """Trains a fresh MLP on rows `idx` of (F, times) with Adam on the GPU. Each epoch uploads a
    seeded permutation of the training rows once; minibatches are consecutive slices of it.
    Returns (mlp, per-step loss). F is (rows, N_FEATURES), times (rows, N_CLASSES): the training
    kernels read minibatches of exactly those widths."""
    F = np.asarray(F, np.float32)
    times = np.asarray(times)
    if F.ndim != 2 or F.shape[1] != N_FEATURES or times.ndim != 2 or times.shape != (F.shape[0], N_CLASSES):
        raise ValueError(f"train: features must be (rows, {N_FEATURES}) and times (rows, {N_CLASSES}); "
                         f"got {F.shape} and {times.shape}")
    F = F[idx]
    T = soft_targets(times[idx], tau)
    mlp = MLP(seed=seed)
    rng = np.random.default_rng(seed)
    N = len(idx)
    batch = min(batch, N)
    hist = []
    pos = N
    dF = dT = None
    for s in range(steps):
        if pos + batch > N:
            order = rng.permutation(N)
            dF = DeviceArray.from_host(np.ascontiguousarray(F[order]).ravel())
            dT = DeviceArray.from_host(np.ascontiguousarray(T[order]).ravel())
            pos = 0
        hist.append(mlp.train_step(dF.ptr + 4 * pos * N_FEATURES, dT.ptr + 4 * pos * N_CLASSES, batch, lr))
        pos += batch
        if log_every and (s + 1) % log_every == 0:
            print(f"step {s + 1}: loss {np.mean(hist[-log_every:]):.4f}", file=sys.stderr, flush=True)
    return mlp, np.array(hist)


def softmax(logits):
    z = np.asarray(logits, np.float64)
    z = np.exp(z - z.max(axis=-1, keepdims=True))
    return z / z.sum(axis=-1, keepdims=True)


def tuning_candidates(probs, min_prob: float = TUNE_MIN_PROB, max_top: int = TUNE_MAX) -> list[int]:
    """Classes to time for one matrix: the most likely first, then every other class of
    probability >= min_prob, at most max_top in all."""
    order = np.argsort(-np.asarray(probs), kind="stable")
    return [int(c) for k, c in enumerate(order[:max(max_top, 1)]) if k == 0 or probs[c] >= min_prob]


def evaluate(chosen, times):
    """Regret of per-matrix choices against the exhaustive best and the default schedule."""
    t = np.asarray(times, np.float64)
    rows = np.arange(len(t))
    tc = t[rows, np.asarray(chosen)]
    best = np.nanmin(t, axis=1)
    regret = tc / best
    vs_default = t[:, DEFAULT_CLASS] / tc
    return dict(count=len(t), mean_regret=float(regret.mean()), geomean_regret=float(np.exp(np.log(regret).mean())),
                p95_regret=float(np.quantile(regret, 0.95)), max_regret=float(regret.max()),
                exact_best=float(np.mean(regret <= 1.0 + 1e-12)),
                geomean_speedup_vs_default=float(np.exp(np.log(vs_default).mean())),
                total_time_ratio_vs_best=float(tc.sum() / best.sum()),
                total_speedup_vs_default=float(t[:, DEFAULT_CLASS].sum() / tc.sum()))


# ---------------------------------------------------------------------------- selector
class Selector:
    """Chooses a schedule for a matrix; falls back to the default schedule (class 0) when no
    trained weights are available, a feature is not finite, the matrix has no non-zeros or the
    network's probability for its choice is below `min_prob`."""

    def __init__(self, weights_path=WEIGHTS_PATH, min_prob: float = 0.0):
        self.min_prob = min_prob
        self.mlp = None
        p = Path(weights_path) if weights_path is not None else None
        if p is not None and p.exists():
            with np.load(p, allow_pickle=False) as z:
                params = z["params"]
            self.mlp = MLP()
            self.mlp.set_params(params)

    def choose_features(self, f) -> tuple[int, float]:
        f = np.asarray(f, np.float32)
        if self.mlp is None or not np.all(np.isfinite(f)) or f[34] <= 0.0:
            return DEFAULT_CLASS, 0.0
        cls, prob = self.mlp.predict(f[None, :])
        c, p = int(cls[0]), float(prob[0])
        if p < self.min_prob:
            return DEFAULT_CLASS, p
        return c, p

    def probs_features(self, f):
        """The network's class probabilities, or None when it falls back to the default."""
        f = np.asarray(f, np.float32)
        if self.mlp is None or not np.all(np.isfinite(f)) or f[34] <= 0.0:
            return None
        return softmax(self.mlp.forward(f[None, :])[0])

    def rank_features(self, f) -> list[int]:
        """Schedules ordered by the network's preference (the default alone as fallback)."""
        p = self.probs_features(f)
        if p is None:
            return [DEFAULT_CLASS]
        return [int(c) for c in np.argsort(-p, kind="stable")]

    def choose(self, m, n, dcolptr, drowidx, stream=None) -> tuple[int, float, np.ndarray]:
        """The network's schedule for a device CSC matrix (fl_select_schedule: features and
        inference on the GPU), its probability, and the features."""
        nnz = drowidx.count if drowidx is not None else 0
        cls, prob = C.c_int(), C.c_float()
        check(lib.fl_select_schedule(self.mlp.handle if self.mlp is not None else None, m, n, nnz, dcolptr.ptr,
                                     drowidx.ptr if nnz else None, C.byref(cls), C.byref(prob),
                                     stream.ptr if stream is not None else None), "fl_select_schedule")
        c, p = int(cls.value), float(prob.value)
        if p < self.min_prob:
            c = DEFAULT_CLASS
        return c, p, features(m, n, dcolptr, drowidx, stream)

    def choose_refined(self, m, n, dcolptr, drowidx, dval, top: int | None = None, reps: int = 10):
        """Model-guided tuning: the network ranks the schedules and the candidates are timed on
        this matrix (`reps` steps each, after warm-up steps); the fastest is returned with the
        measured times. Candidates: the `top` best-ranked, or by default (top=None) those of
        tuning_candidates (probability >= TUNE_MIN_PROB, at most TUNE_MAX), plus the HYB class
        when the HYB_COLS longest columns hold at least HYB_MIN_SHARE of the entries. One-time
        cost: preparing and timing the candidates."""
        f = features(m, n, dcolptr, drowidx)
        if top is None:
            p = self.probs_features(f)
            cands = [DEFAULT_CLASS] if p is None else tuning_candidates(p)
        else:
            cands = self.rank_features(f)[:max(top, 1)]
        if HYB_CLASS not in cands and heavy_share(n, dcolptr) >= HYB_MIN_SHARE:
            cands = cands + [HYB_CLASS]
        if len(cands) == 1:
            return cands[0], {cands[0]: None}, f
        mats, times = {}, {}
        dx = DeviceArray.from_host(np.ones(n, np.float32))
        dy = DeviceArray(max(m, 1), np.float32)
        for c in cands:
            run = ScheduledSpMV(m, n, dcolptr, drowidx, dval, c, mats)
            run.set_x(dx)
            times[c] = time_schedule(run.step, dy, reps_min=reps, reps_max=reps, rounds=1)
        for M in mats.values():
            M.close()
        best = min(times, key=times.get)
        return best, times, f


# ---------------------------------------------------------------- schedule execution
class ScheduledSpMV:
    """Runs y = A x in FP32 with one schedule on a resident matrix. The device CSC arrays are
    only read during construction."""

    def __init__(self, m, n, dcolptr, drowidx, dval, cls: int, mats=None):
        self.cls = cls
        self.sched = s = SCHEDULES[cls]
        self.m, self.n = m, n
        key = s.width
        if mats is not None and key in mats:
            M = mats[key]
        else:
            M = FastLoadMatrix(m, n, dcolptr, drowidx, dval, width=s.width)
            if mats is not None:
                mats[key] = M
        self.M = M
        if s.kernel == KERNEL_ROW:
            M.prepare_rows()
        else:
            M.prepare(VAL_F32)
            if s.kernel in (KERNEL_WIN, KERNEL_WINR, KERNEL_PB):
                M.prepare_schedule(s.kernel, s.batch_tiles, max(s.band_rows, 1024))
            elif s.kernel == KERNEL_HYB:
                M.prepare_hyb(min(n, HYB_COLS), s.batch_tiles, s.band_rows)
        self.dsx = None

    def set_x(self, dx: DeviceArray):
        """x in the original column order (float32)."""
        if self.sched.kernel == KERNEL_ROW:
            self.dsx = dx
        else:
            if self.dsx is None or self.dsx is dx:
                self.dsx = DeviceArray(self.n, np.float32)
            self.M.permute_x(dx, self.dsx)

    def step(self, dy: DeviceArray, stream=None):
        """y = A x (float32 y of m entries)."""
        if self.sched.kernel == KERNEL_ROW:
            self.M.spmm(self.dsx, dy, 1, VAL_F32, accumulate=False, stream=stream)
        else:
            self.M.spmv(self.dsx, dy, self.sched.kernel, VAL_F32, stream=stream, assign=True)


# ------------------------------------------------------------------ sweep population
POPULATION_FAMILIES = ("zipf", "banded", "blockdiag", "uniform", "zipf_dense", "stencil")
NNZ_CAP = 72_000_000
N_MIN, N_MAX = 10_000, 8_388_608


def _zipf_mean(a, kmax):
    k = np.arange(1, int(kmax) + 1, dtype=np.float64)
    w = k ** -a
    return float((k * w).sum() / w.sum())


def population_spec(i: int, base_seed: int = 20240):
    """Matrix i of the selector population: (family name, gen family, m, n, params, seed).

    Family i % 6; columns log-uniform in [10k, 8M], square; family parameters:
      zipf        exponent U(1.5, 2.5), max column length min(m, 50,000)
      banded      width log-uniform in [5, 200]
      blockdiag   block size log-uniform in [32, 1024], entries per column log-uniform in
                  [2, block]
      uniform     density log-uniform in [1e-6, 1e-3]
      zipf_dense  zipf as above plus 0.1% columns of log-uniform length in [5k, 200k] (<= m)
      stencil     27-point stencil on the cube of the nearest side (27 diagonals: multi-band)
    The expected non-zero count is kept below 72M, which bounds a sweep while covering the
    67M-entry power-law workload: columns are reduced (not below 10k) for zipf, banded, blockdiag
    and the stencil; for uniform the density is reduced instead, so every matrix keeps between
    10k and 8M columns."""
    rng = np.random.default_rng([base_seed, i])
    fam = POPULATION_FAMILIES[i % len(POPULATION_FAMILIES)]
    n = int(round(math.exp(rng.uniform(math.log(N_MIN), math.log(N_MAX)))))
    seed = base_seed * 100_000 + i
    if fam in ("zipf", "zipf_dense"):
        a = float(rng.uniform(1.5, 2.5))
        kmax = 50_000
        dense_len = float(round(math.exp(rng.uniform(math.log(5e3), math.log(2e5)))))
        mean = _zipf_mean(a, min(kmax, n))
        if fam == "zipf_dense":
            mean = 0.999 * mean + 0.001 * min(dense_len, n)
        n = max(N_MIN, min(n, int(NNZ_CAP / mean)))
        if fam == "zipf":
            return fam, GEN_ZIPF, n, n, (a, kmax), seed
        return fam, GEN_ZIPF_DENSE, n, n, (a, kmax, 0.001, dense_len), seed
    if fam == "banded":
        w = int(round(math.exp(rng.uniform(math.log(5), math.log(200)))))
        n = max(N_MIN, min(n, NNZ_CAP // (w + 1)))
        return fam, GEN_BANDED, n, n, (w,), seed
    if fam == "blockdiag":
        b = int(round(math.exp(rng.uniform(math.log(32), math.log(1024)))))
        k = int(round(math.exp(rng.uniform(math.log(2), math.log(b)))))
        n = max(N_MIN, min(n, NNZ_CAP // k))
        return fam, GEN_BLOCKDIAG, n, n, (b, k), seed
    if fam == "stencil":
        side = int(round(n ** (1.0 / 3.0)))
        side = max(22, min(side, int((NNZ_CAP / 27) ** (1.0 / 3.0))))  # 22^3 = 10,648 columns
        return fam, GEN_STENCIL27, side ** 3, side ** 3, (side,), seed
    d = float(math.exp(rng.uniform(math.log(1e-6), math.log(1e-3))))
    d = min(d, NNZ_CAP / float(n) / float(n))
    return fam, GEN_UNIFORM, n, n, (d,), seed


def population_split(count: int = 2000):
    """Held-out set: every fifth group of five consecutive matrices (400 of 2,000, 66 or 67 of
    each of the six families); the rest train."""
    idx = np.arange(count)
    test = (idx // 5) % 5 == 4
    return idx[~test], idx[test]


# --------------------------------------------------------------------------- sweeping
def time_schedule(run, dy, reps_min=5, reps_max=60, target_ms=3.0, rounds=3):
    """Mean step time (ms) of the fastest of `rounds` timed batches; a batch is `reps` steps
    between two events, reps sized so a batch lasts about target_ms."""
    for _ in range(2):
        run(dy)
    a, b = Event(), Event()
    a.record()
    run(dy)
    b.record()
    synchronize()
    t1 = max(a.elapsed_ms(b), 1e-3)
    reps = int(min(reps_max, max(reps_min, target_ms / t1)))
    best = math.inf
    for _ in range(rounds):
        a.record()
        for _ in range(reps):
            run(dy)
        b.record()
        synchronize()
        best = min(best, a.elapsed_ms(b) / reps)
    return best


def sweep_one(i: int, base_seed: int = 20240, rounds: int = 3):
    """Features, per-schedule step times (ms) and per-schedule relative L2 difference to the
    default schedule's y for matrix i of the population."""
    fam, gfam, m, n, params, seed = population_spec(i, base_seed)
    t0 = time.time()
    m, n, cp, ri, v = gen_matrix(gfam, seed, m, n, params)
    nnz = int(cp[n])
    tgen = time.time() - t0
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    del cp, ri, v
    f = features(m, n, dcp, dri)
    x = np.random.default_rng(seed).standard_normal(n).astype(np.float32)
    dx = DeviceArray.from_host(x)
    dy = DeviceArray(max(m, 1), np.float32)
    mats = {}
    times = np.full(N_CLASSES, np.inf)
    errs = np.full(N_CLASSES, np.nan)
    y0 = None
    for c in range(N_CLASSES):
        r = ScheduledSpMV(m, n, dcp, dri, dv, c, mats)
        r.set_x(dx)
        times[c] = time_schedule(r.step, dy, rounds=rounds)
        r.step(dy)
        synchronize()
        y = dy.to_host().astype(np.float64)
        if y0 is None:
            y0 = y
            nrm = float(np.linalg.norm(y0)) or 1.0
        errs[c] = float(np.linalg.norm(y - y0)) / nrm
        if r.dsx is not dx:
            r.dsx.free()
    for M in mats.values():
        M.close()
    meta = dict(family=fam, m=m, n=n, nnz=nnz, params=list(params), seed=seed, gen_s=round(tgen, 3),
                total_s=round(time.time() - t0, 3))
    return f, times, errs, meta


def sweep(start: int, stop: int, out: str, base_seed: int = 20240, save_every: int = 20):
    import json
    F, Tm, E, metas, ids = [], [], [], [], []

    def save():
        np.savez_compressed(out, index=np.array(ids, np.int32), features=np.array(F, np.float32).reshape(-1, N_FEATURES),
                            times=np.array(Tm, np.float64).reshape(-1, N_CLASSES),
                            errs=np.array(E, np.float64).reshape(-1, N_CLASSES),
                            meta=np.array([json.dumps(mm) for mm in metas]))

    for i in range(start, stop):
        f, t, e, meta = sweep_one(i, base_seed)
        F.append(f), Tm.append(t), E.append(e), metas.append(meta), ids.append(i)
        b = int(np.argmin(t))
        print(f"[{i}] {meta['family']:10s} n={meta['n']:8d} nnz={meta['nnz']:9d} best={SCHEDULES[b].name:11s} "
              f"{t[b]:.4f} ms default {t[DEFAULT_CLASS]:.4f} ms maxdiff {np.nanmax(e):.1e} ({meta['total_s']:.1f}s)",
              flush=True)
        if (i - start + 1) % save_every == 0:
            save()
    save()


def load_dataset(paths):
    """Concatenates sweep files (sorted by matrix index, later files win on duplicates)."""
    import json
    rows = {}
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            for k, i in enumerate(z["index"]):
                rows[int(i)] = (z["features"][k], z["times"][k], z["errs"][k], json.loads(str(z["meta"][k])))
    ids = np.array(sorted(rows), np.int32)
    F = np.array([rows[i][0] for i in ids], np.float32)
    T = np.array([rows[i][1] for i in ids], np.float64)
    E = np.array([rows[i][2] for i in ids], np.float64)
    M = [rows[i][3] for i in ids]
    return ids, F, T, E, M


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep", help="time every schedule on matrices [start, stop) of the population")
    s.add_argument("--start", type=int, required=True)
    s.add_argument("--stop", type=int, required=True)
    s.add_argument("--out", required=True)
    t = sub.add_parser("train", help="train on the sweep files, evaluate on the held-out split, save weights")
    t.add_argument("--data", nargs="+", default=[str(DATASET_PATH)])
    t.add_argument("--out", default=str(WEIGHTS_PATH))
    t.add_argument("--steps", type=int, default=5000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--tau", type=float, default=0.05)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--tune", action="store_true",
                   help="pick tau and steps on a validation part of the training matrices first")
    t.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "sweep":
        sweep(a.start, a.stop, a.out)
        return 0
    ids, F, T, E, M = load_dataset(a.data)
    tr_all, te_all = population_split(2000)
    tr = np.flatnonzero(np.isin(ids, tr_all))
    te = np.flatnonzero(np.isin(ids, te_all))
    rep = dict(train=len(tr), test=len(te))
    tau, steps = a.tau, a.steps
    if a.tune:
        # validation: every fifth group of five among the training matrices (never the test set)
        val = tr[(ids[tr] // 5) % 5 == 3]
        fit = tr[(ids[tr] // 5) % 5 != 3]
        grid = []
        for g_tau in (0.02, 0.05, 0.1):
            for g_steps in (2000, 5000):
                mlp, _ = train(F, T, fit, steps=g_steps, lr=a.lr, tau=g_tau, seed=a.seed)
                r = evaluate(mlp.predict(F[val])[0], T[val])
                grid.append(dict(tau=g_tau, steps=g_steps, val_geomean_regret=r["geomean_regret"],
                                 val_mean_regret=r["mean_regret"]))
                print(json.dumps(grid[-1]), file=sys.stderr, flush=True)
                mlp.close()
        pick = min(grid, key=lambda d: d["val_geomean_regret"])
        tau, steps = pick["tau"], pick["steps"]
        rep["tuning"] = dict(fit=len(fit), val=len(val), grid=grid, picked=dict(tau=tau, steps=steps))
    mlp, hist = train(F, T, tr, steps=steps, lr=a.lr, tau=tau, seed=a.seed, log_every=1000)
    cls, prob = mlp.predict(F[te])
    rep.update(tau=tau, steps=steps, lr=a.lr, final_loss=float(hist[-50:].mean()),
               learned=evaluate(cls, T[te]), default=evaluate(np.full(len(te), DEFAULT_CLASS), T[te]))
    fixed = np.exp(np.log(T[tr] / T[tr].min(axis=1, keepdims=True)).mean(axis=0))
    bf = int(np.argmin(fixed))
    rep["best_fixed"] = dict(cls=bf, name=SCHEDULES[bf].name, **evaluate(np.full(len(te), bf), T[te]))
    logits = mlp.forward(F[te])
    for k in (2, 3):
        topk = np.argsort(-logits, axis=1, kind="stable")[:, :k]
        pick = topk[np.arange(len(te)), np.argmin(np.take_along_axis(T[te], topk, axis=1), axis=1)]
        rep[f"learned_top{k}_timed"] = evaluate(pick, T[te])
    cands = [tuning_candidates(p) for p in softmax(logits)]
    pick = [c[int(np.argmin(T[te][i, c]))] for i, c in enumerate(cands)]
    rep["learned_adaptive_timed"] = dict(min_prob=TUNE_MIN_PROB, max_top=TUNE_MAX,
                                         mean_candidates=float(np.mean([len(c) for c in cands])),
                                         **evaluate(pick, T[te]))
    rep["chosen_hist"] = {SCHEDULES[c].name: int((cls == c).sum()) for c in range(N_CLASSES)}
    rep["best_hist"] = {SCHEDULES[c].name: int((T[te].argmin(1) == c).sum()) for c in range(N_CLASSES)}
    print(json.dumps(rep, indent=1))
    np.savez_compressed(a.out, params=mlp.get_params(), dims=np.array(mlp.dims, np.int32),
                        tau=np.float32(tau), steps=np.int32(steps), lr=np.float32(a.lr), seed=np.int32(a.seed))
    if a.report:
        Path(a.report).write_text(json.dumps(rep, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
