"""CPU checks of the learned schedule selector: the selector oracle is pinned (brute-force loops,
a closed form, finite differences, bit patterns), and the host logic (population, split, soft
targets, regret, fallback) behaves as documented. When the committed sweep and weights exist,
their consistency and the trained network's gain are checked with the oracle's forward pass."""
import json
import math
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import selector_oracle as so  # noqa: E402

import fastload_amd as fl  # noqa: E402
from fastload_amd import selector as sel  # noqa: E402


# ---------------------------------------------------------------------------- oracle pins
def test_bf16_bit_patterns():
    e = 2.0 ** -8
    vals = np.array([1.0, 1 + e, 1 + 3 * e, 2 - e, 1 + e + 2.0 ** -20, -(1 + 3 * e), 0.0, -0.0, 3.4e38, 1e-40],
                    np.float32)
    # 3.4e38 lies above the largest bfloat16 (0x7F7F) by more than half an ulp: infinity;
    # 1e-40 is the subnormal 0x000116C2, whose low half is below the tie: 0x00010000
    want = np.array([1.0, 1.0, 1 + 4 * e, 2.0, 1 + 2 * e, -(1 + 4 * e), 0.0, -0.0, np.inf,
                     np.array([0x00010000], np.uint32).view(np.float32)[0]], np.float32)
    got = so.bf16(vals)
    for g, w, v in zip(got, want, vals):
        assert g == w, (v, g, w)
    assert np.signbit(got[7])
    assert np.isnan(so.bf16(np.array([np.nan], np.float32)))[0]


def brute_features(m, n, cp, ri):
    hist = [0] * 32
    spans = []
    for c in range(n):
        a, b = int(cp[c]), int(cp[c + 1])
        L = b - a
        hist[0 if L == 0 else min(31, int(math.floor(math.log2(L))) + 1)] += 1
        if L:
            spans.append((int(ri[b - 1]) - int(ri[a])) / m)
    rc = [0] * m
    for r in ri[:int(cp[n])]:
        rc[int(r)] += 1
    nnz = int(cp[n])
    mean = nnz / m
    var = sum((v - mean) ** 2 for v in rc) / m
    f = [h / n for h in hist]
    f += [math.log2(m + 1) / 24, math.log2(n + 1) / 24, math.log2(nnz + 1) / 32, math.log2(mean + 1) / 8,
          math.sqrt(var) / mean if mean else 0.0, math.log2(max(rc) / mean + 1) / 8 if mean else 0.0,
          sum(v == 0 for v in rc) / m, sum(spans) / len(spans) if spans else 0.0]
    return np.array(f)


@pytest.mark.parametrize("case", range(6))
def test_feature_oracle_matches_brute_force(case):
    rng = np.random.default_rng(case)
    m, n = int(rng.integers(5, 300)), int(rng.integers(5, 300))
    dens = [0.01, 0.05, 0.3, 0.0, 0.9, 0.002][case]
    A = rng.random((m, n)) < dens
    if case == 4:
        A[:, :3] = True  # dense columns
    cp = np.concatenate([[0], np.cumsum(A.sum(axis=0))]).astype(np.int32)
    ri = np.concatenate([np.flatnonzero(A[:, c]) for c in range(n)] + [np.zeros(0, np.int64)]).astype(np.int32)
    got = so.features(m, n, cp, ri)
    want = brute_features(m, n, cp, ri)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-7), np.flatnonzero(~np.isclose(got, want))


def test_feature_oracle_band_closed_form():
    n = 1000
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_BANDED, 1, n, n, (9,))
    f = so.features(m, n, cp, ri)
    # columns 0..3 and n-4..n-1 hold 5,6,7,8 rows; the rest 9: bin 3 = [4,8), bin 4 = [8,16)
    assert f[3] == np.float32(6 / n) and f[4] == np.float32(994 / n) and f[:32].sum() == pytest.approx(1.0)
    nnz = 992 * 9 + 2 * (5 + 6 + 7 + 8)
    assert int(cp[n]) == nnz
    assert f[34] == np.float32(math.log2(nnz + 1) / 32)
    span = (992 * 8 + 2 * (4 + 5 + 6 + 7)) / n / n
    assert f[39] == pytest.approx(span, rel=1e-6)
    assert f[38] == 0.0
    mean = nnz / n
    var = (992 * 81 + 2 * (25 + 36 + 49 + 64)) / n - mean * mean
    assert f[36] == pytest.approx(math.sqrt(var) / mean, rel=1e-6)


def test_fp64_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = (40, 64, 64, 12)
    P = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(3))
    params = rng.standard_normal(P) * 0.3
    X = rng.random((9, 40))
    T = rng.random((9, 12))
    T /= T.sum(axis=1, keepdims=True)
    g, loss = so.mlp_backward(params, dims, X, T, rounding=False)
    idx = rng.choice(P, 200, replace=False)
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (so.mlp_forward(p1, dims, X, T)[2] - so.mlp_forward(p2, dims, X, T)[2]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # the loss is the mean cross-entropy: a uniform prediction gives log(12)
    zero = np.zeros(P)
    assert so.mlp_forward(zero, dims, X, T)[2] == pytest.approx(math.log(12))


# ----------------------------------------------------------------------------- host logic
def test_schedule_table():
    assert sel.N_CLASSES == 13 and len({s.name for s in sel.SCHEDULES}) == 13
    assert sel.SCHEDULES[sel.HYB_CLASS].kernel == fl.KERNEL_HYB
    d = sel.SCHEDULES[sel.DEFAULT_CLASS]
    assert (d.width, d.kernel) == (32, fl.KERNEL_DROPIN)  # the reference schedule
    for s in sel.SCHEDULES:
        if s.kernel in (fl.KERNEL_WIN, fl.KERNEL_WINR, fl.KERNEL_PB):
            assert s.batch_tiles in (64, 128, 256) and s.width == 64
        if s.kernel in (fl.KERNEL_PB, fl.KERNEL_HYB):
            assert 1024 <= s.band_rows <= 16384


def test_population_and_split():
    specs = [sel.population_spec(i) for i in range(2000)]
    assert specs[17] == sel.population_spec(17)
    fams = [s[0] for s in specs]
    nf = len(sel.POPULATION_FAMILIES)
    assert {f: fams.count(f) for f in set(fams)} == {f: 2000 // nf + (k < 2000 % nf)
                                                     for k, f in enumerate(sel.POPULATION_FAMILIES)}
    for fam, g, m, n, p, seed in specs:
        assert m == n and sel.N_MIN <= n <= sel.N_MAX
        if fam == "zipf":
            exp_nnz = n * sel._zipf_mean(p[0], min(p[1], n))
        elif fam == "zipf_dense":
            exp_nnz = n * (0.999 * sel._zipf_mean(p[0], min(p[1], n)) + 0.001 * min(p[3], n))
        elif fam == "banded":
            exp_nnz = n * p[0]
        elif fam == "blockdiag":
            exp_nnz = n * p[1]
            assert 32 <= p[0] <= 1024 and 2 <= p[1] <= p[0]
        elif fam == "stencil":
            assert n == p[0] ** 3 and g == fl.GEN_STENCIL27
            exp_nnz = 27 * n
        else:
            exp_nnz = n * n * p[0]
            assert 1e-6 <= p[0] <= 1e-3
        assert exp_nnz <= sel.NNZ_CAP * 1.001
    tr, te = sel.population_split(2000)
    assert len(tr) == 1600 and len(te) == 400 and not set(tr) & set(te)
    assert all(sum(1 for i in te if fams[i] == f) in (66, 67) for f in sel.POPULATION_FAMILIES)


def test_soft_targets_and_regret():
    t = np.array([[1.0, 1.05, 2.0, np.inf], [3.0, 1.5, 1.5, 6.0]])
    T = sel.soft_targets(t, tau=0.05)
    assert np.allclose(T.sum(axis=1), 1.0) and T[0, 3] == 0.0
    assert T[0, 0] / T[0, 1] == pytest.approx(math.exp(math.log(1.05) / 0.05), rel=1e-5)
    assert T[1, 1] == pytest.approx(T[1, 2])
    times = np.array([[2.0, 1.0, 4.0], [1.0, 3.0, 2.0]])
    r = sel.evaluate([1, 2], times)
    assert r["mean_regret"] == pytest.approx(1.5) and r["exact_best"] == 0.5
    assert r["geomean_speedup_vs_default"] == pytest.approx(math.sqrt(2 * 0.5))


def test_selector_fallback_without_weights():
    s = sel.Selector(weights_path=None)
    f = np.zeros(40, np.float32)
    f[34] = 0.5
    assert s.choose_features(f) == (sel.DEFAULT_CLASS, 0.0)
    f[3] = np.nan
    assert s.choose_features(f)[0] == sel.DEFAULT_CLASS


# ------------------------------------------------------------- committed sweep and weights
needs_data = pytest.mark.skipif(not (sel.DATASET_PATH.exists() and sel.WEIGHTS_PATH.exists()),
                                reason="selector sweep/weights not committed")


@needs_data
def test_dataset_is_complete_and_consistent():
    ids, F, T, E, M = sel.load_dataset([sel.DATASET_PATH])
    assert np.array_equal(ids, np.arange(2000))
    assert np.all(np.isfinite(T)) and np.all(T > 0)
    assert np.nanmax(E) <= 1e-4  # every schedule's FP32 result agrees with the default's
    for i in (0, 1, 3, 5, 8):  # small matrices: regenerate on the host and recompute the features
        fam, g, m, n, p, seed = sel.population_spec(i)
        assert M[i]["family"] == fam and M[i]["n"] == n
        if M[i]["nnz"] > 3_000_000:
            continue
        m, n, cp, ri, v = fl.gen_matrix(g, seed, m, n, p)
        assert int(cp[n]) == M[i]["nnz"]
        assert np.allclose(F[i], so.features(m, n, cp, ri), rtol=1e-5, atol=1e-6)


@needs_data
def test_trained_network_gains_over_default_and_fixed_schedule():
    """The committed network, evaluated with the oracle's forward pass (BF16 rounding as on the
    GPU) on the 400 held-out matrices, must beat the default schedule and the best single fixed
    schedule chosen on the training matrices."""
    ids, F, T, E, M = sel.load_dataset([sel.DATASET_PATH])
    tr, te = sel.population_split(2000)
    with np.load(sel.WEIGHTS_PATH, allow_pickle=False) as z:
        params, dims = z["params"], tuple(int(d) for d in z["dims"])
    logits = so.mlp_forward(params, dims, F[te], rounding=True)[0]
    chosen = logits.argmax(axis=1)
    learned = sel.evaluate(chosen, T[te])
    default = sel.evaluate(np.zeros(len(te), int), T[te])
    fixed_score = np.exp(np.log(T[tr] / T[tr].min(axis=1, keepdims=True)).mean(axis=0))
    bf = int(np.argmin(fixed_score))
    fixed = sel.evaluate(np.full(len(te), bf), T[te])
    print(json.dumps(dict(learned=learned, default=default, fixed=fixed), indent=1))
    assert learned["geomean_speedup_vs_default"] > 1.5
    assert learned["geomean_regret"] < fixed["geomean_regret"]
    assert learned["geomean_regret"] < 1.10
    # model-guided tuning: timing the three best-ranked schedules can only lower the regret
    top3 = np.argsort(-logits, axis=1, kind="stable")[:, :3]
    pick = top3[np.arange(len(te)), np.argmin(np.take_along_axis(T[te], top3, axis=1), axis=1)]
    tuned = sel.evaluate(pick, T[te])
    assert tuned["geomean_regret"] <= learned["geomean_regret"] and tuned["exact_best"] >= learned["exact_best"]
    # the default tuning rule (every schedule of probability >= 2 %, at most 5) is at least as
    # good as the top three and times no more schedules than they do on average
    cands = [sel.tuning_candidates(p) for p in sel.softmax(logits)]
    assert all(1 <= len(c) <= sel.TUNE_MAX and c[0] == int(np.argmax(lg)) for c, lg in zip(cands, logits))
    pick = [c[int(np.argmin(T[te][i, c]))] for i, c in enumerate(cands)]
    adaptive = sel.evaluate(pick, T[te])
    assert adaptive["geomean_regret"] <= tuned["geomean_regret"] + 1e-9
    assert np.mean([len(c) for c in cands]) <= 3.0


def test_train_rejects_mismatched_widths():
    """The GPU training kernels read minibatches N_FEATURES and N_CLASSES wide: other widths are
    refused before anything reaches the device."""
    F = np.zeros((8, sel.N_FEATURES), np.float32)
    with pytest.raises(ValueError):
        sel.train(F, np.ones((8, sel.N_CLASSES - 1)), np.arange(8))
    with pytest.raises(ValueError):
        sel.train(F[:, :-1], np.ones((8, sel.N_CLASSES)), np.arange(8))
    with pytest.raises(ValueError):
        sel.train(F, np.ones((7, sel.N_CLASSES)), np.arange(7))


def test_tuning_candidates_rule():
    """Ordered by probability; the first always, the rest only at >= TUNE_MIN_PROB; at most
    TUNE_MAX."""
    p = np.array([0.01, 0.5, 0.049, 0.2, 0.05, 0.06, 0.07, 0.051, 0.0, 0.0, 0.0, 0.0])
    assert sel.tuning_candidates(p, min_prob=0.05) == [1, 3, 6, 5, 7]
    assert sel.tuning_candidates(p, min_prob=0.05, max_top=2) == [1, 3]
    assert sel.tuning_candidates(p) == [1, 3, 6, 5, 7] and sel.tuning_candidates(p, max_top=6) == [1, 3, 6, 5, 7, 4]
    assert sel.tuning_candidates(np.full(12, 1 / 12), min_prob=0.5) == [0]
    q = sel.softmax(np.array([[1.0, 2.0, 3.0]]))
    assert np.allclose(q.sum(), 1.0) and np.argmax(q) == 2


def test_selector_entry_points_reject_bad_arguments():
    """Host-side validation of the selector's C entry points (no device work is started)."""
    import ctypes as C
    lib = fl.lib
    out = np.zeros(40, np.float32)
    assert lib.fl_features(-1, 5, 0, None, None, out.ctypes.data, None) == -8  # FL_ERR_INVALID
    assert lib.fl_features(5, 5, 3, None, None, out.ctypes.data, None) == -8
    h = C.c_void_p()
    assert lib.fl_mlp_create(C.byref(h), 0, 64, 64, 12, 0) == -8
    assert lib.fl_mlp_create(None, 40, 64, 64, 12, 0) == -8
    assert lib.fl_mlp_param_count(None) == 0
    assert lib.fl_mlp_forward(None, None, 1, None, None) == -8
    assert lib.fl_mlp_predict(None, None, 1, None, None, None) == -8
    assert lib.fl_mlp_gradients(None, None, None, 1, None, None, None) == -8
    assert lib.fl_mlp_train_step(None, None, None, 1, C.c_float(0.1), None, None) == -8
    assert b"fl_mlp_train_step" in lib.fl_last_error()


def test_schedule_table_comes_from_the_library():
    import ctypes as C
    names = [s.name for s in sel.SCHEDULES]
    assert names == ["dropin_w32", "dropin_w64", "winr64", "win64", "win128", "win256", "pb128_r4k", "pb128_r16k",
                     "pb256_r4k", "pb256_r8k", "pb256_r16k", "row", "hyb32k_pb128_r16k"]
    w = C.c_int()
    assert fl.lib.fl_schedule_info(13, C.byref(w), None, None, None) == -8
    assert fl.lib.fl_schedule_info(-1, None, None, None, None) == -8
    assert fl.lib.fl_select_schedule(None, 5, 5, 0, None, None, None, None, None) == -8
