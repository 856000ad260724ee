"""GPU tests of the learned schedule selector's kernels.

Pins: fl_features against the selector oracle (histogram exact, statistics to 1e-5); the MLP
forward pass and gradients against the oracle's FP64 computation with the same BF16 rounding
points (a transposed or mis-laid-out MFMA operand changes every output); the BF16 effect itself
against the exact FP64 gradients; one Adam step against numpy; training on a learnable problem;
the committed network's gain on the held-out sweep matrices; and an end-to-end selected SpMV
against the FP64 oracle."""
import sys

import numpy as np
import pytest

import oracle_bind as ob
from conftest import ROOT
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray
from fastload_amd import selector as sel

sys.path.insert(0, str(ROOT / "oracle"))
import selector_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu

DIMS = (40, 64, 64, 12)


def feature_cases():
    yield fl.gen_matrix(fl.GEN_ZIPF, 3, 5000, 7000, (1.8, 2000))
    yield fl.gen_matrix(fl.GEN_BANDED, 4, 3000, 3000, (37,))
    yield fl.gen_matrix(fl.GEN_BLOCKDIAG, 5, 4000, 4000, (100, 7))
    yield fl.gen_matrix(fl.GEN_UNIFORM, 6, 20000, 9000, (2e-4,))
    yield fl.gen_matrix(fl.GEN_ZIPF_DENSE, 7, 8000, 8000, (2.0, 5000, 0.01, 6000))
    yield fl.gen_matrix(fl.GEN_STENCIL27, 8, 0, 0, (12,))
    yield fl.gen_matrix(fl.GEN_UNIFORM, 9, 300, 50, (1e-3,))  # many empty columns and rows


@pytest.mark.parametrize("case", range(7))
def test_features_match_oracle(gpu, case):
    m, n, cp, ri, v = list(feature_cases())[case]
    f = sel.features(m, n, DeviceArray.from_host(cp), DeviceArray.from_host(ri) if len(ri) else None)
    want = so.features(m, n, cp, ri)
    assert np.array_equal(f[:32], want[:32])
    assert np.allclose(f[32:], want[32:], rtol=1e-5, atol=1e-7), (f[32:], want[32:])


def test_features_empty_matrix(gpu):
    cp = np.zeros(11, np.int32)
    f = sel.features(5, 10, DeviceArray.from_host(cp), None)
    assert f[0] == 1.0 and f[34] == 0.0 and np.all(np.isfinite(f))


def rand_problem(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((B, 40)).astype(np.float32)
    T = rng.random((B, 12)).astype(np.float32) ** 4
    T /= T.sum(axis=1, keepdims=True)
    return X, T


def test_params_roundtrip_and_forward(gpu):
    mlp = sel.MLP(DIMS, seed=1)
    p = mlp.get_params()
    assert p.size == sum(DIMS[l] * DIMS[l + 1] + DIMS[l + 1] for l in range(3)) == mlp.count
    rng = np.random.default_rng(2)
    p2 = (rng.standard_normal(p.size) * 0.2).astype(np.float32)
    mlp.set_params(p2)
    assert np.array_equal(mlp.get_params(), p2)
    for B in (1, 7, 64, 200):
        X, _ = rand_problem(B, B)
        L = mlp.forward(X)
        want = so.mlp_forward(p2, DIMS, X, rounding=True)[0]
        scale = np.abs(want).max()
        assert np.abs(L - want).max() <= 1e-4 * scale, B
        cls, prob = mlp.predict(X)
        margin = np.sort(want, axis=1)
        clear = margin[:, -1] - margin[:, -2] > 1e-3 * scale
        assert np.array_equal(cls[clear], want.argmax(axis=1)[clear])
        sm = np.exp(want - want.max(axis=1, keepdims=True))
        assert np.allclose(prob, (sm / sm.sum(axis=1, keepdims=True)).max(axis=1), rtol=1e-3)


@pytest.mark.parametrize("B", [16, 100, 256])
def test_gradients_match_cpu_fp64_backward(gpu, B):
    mlp = sel.MLP(DIMS, seed=3)
    p = mlp.get_params()
    X, T = rand_problem(B, 10 + B)
    g, loss = mlp.gradients(X, T)
    want, wloss = so.mlp_backward(p, DIMS, X, T, rounding=True)
    exact, _ = so.mlp_backward(p, DIMS, X, T, rounding=False)
    assert loss == pytest.approx(wloss, rel=1e-4)
    o = 0
    for l in range(3):
        for k in (DIMS[l] * DIMS[l + 1], DIMS[l + 1]):
            a, b, e = g[o:o + k], want[o:o + k], exact[o:o + k]
            nb = np.linalg.norm(b)
            assert np.linalg.norm(a - b) <= 2e-3 * nb, (l, k, np.linalg.norm(a - b) / nb)
            # BF16 operands move the gradient by about their relative precision, no more
            assert np.linalg.norm(a - e) <= 3e-2 * np.linalg.norm(e), (l, k)
            o += k


def test_adam_step_matches_numpy(gpu):
    mlp = sel.MLP(DIMS, seed=4)
    p0 = mlp.get_params().astype(np.float64)
    X, T = rand_problem(64, 5)
    g, _ = mlp.gradients(X, T)
    dX, dT = DeviceArray.from_host(X.ravel()), DeviceArray.from_host(T.ravel())
    lr = 1e-2
    mlp.train_step(dX.ptr, dT.ptr, 64, lr)
    p1 = mlp.get_params()
    m = 0.1 * g
    v = 0.001 * g.astype(np.float64) ** 2
    want = p0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(p1, want, rtol=1e-5, atol=1e-6)


def test_training_learns_a_separable_problem(gpu):
    rng = np.random.default_rng(6)
    X = rng.random((1024, 40)).astype(np.float32)
    Wt = rng.standard_normal((40, sel.N_CLASSES))
    lab = (X @ Wt).argmax(axis=1)
    times = np.where(np.eye(sel.N_CLASSES)[lab] > 0, 1.0, 1.5)
    idx = np.arange(1024)
    mlp, hist = sel.train(X, times, idx[:768], steps=1500, batch=128, lr=3e-3, seed=0)
    assert hist[-50:].mean() < 0.5 * hist[:10].mean()
    cls, _ = mlp.predict(X[768:])
    assert np.mean(cls == lab[768:]) > 0.85


needs_data = pytest.mark.skipif(not (sel.DATASET_PATH.exists() and sel.WEIGHTS_PATH.exists()),
                                reason="selector sweep/weights not committed")


@needs_data
def test_selector_gain_on_held_out_matrices(gpu):
    """The committed network run by the GPU kernels on the 400 held-out sweep matrices: it beats
    the default schedule and the best fixed schedule; a network retrained here from the sweep
    with the GPU training kernels does too."""
    ids, F, T, E, M = sel.load_dataset([sel.DATASET_PATH])
    tr, te = sel.population_split(2000)
    s = sel.Selector()
    chosen = np.array([s.choose_features(F[i])[0] for i in te])
    learned = sel.evaluate(chosen, T[te])
    fixed_score = np.exp(np.log(T[tr] / T[tr].min(axis=1, keepdims=True)).mean(axis=0))
    fixed = sel.evaluate(np.full(len(te), int(np.argmin(fixed_score))), T[te])
    print(learned, fixed)
    assert learned["geomean_speedup_vs_default"] > 1.5
    assert learned["geomean_regret"] < fixed["geomean_regret"]
    mlp, hist = sel.train(F, T, tr, steps=2000, seed=1)
    cls, _ = mlp.predict(F[te])
    re = sel.evaluate(cls, T[te])
    assert re["geomean_speedup_vs_default"] > 1.5 and re["geomean_regret"] < fixed["geomean_regret"]


@pytest.mark.parametrize("cls", range(sel.N_CLASSES))
def test_every_schedule_class_runs_correctly(gpu, cls):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 12, 30000, 30000, (1.8, 3000))
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    run = sel.ScheduledSpMV(m, n, dcp, dri, dv, cls)
    x = np.random.default_rng(cls).standard_normal(n)
    dx = DeviceArray.from_host(x.astype(np.float32))
    run.set_x(dx)
    dy = DeviceArray(m, np.float32)
    dy.copy_from(DeviceArray.from_host(np.full(m, 5.0, np.float32)))  # step must overwrite y
    run.step(dy)
    yo, S = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x.astype(np.float32))
    assert rel_err(dy.to_host().astype(np.float64), yo, S) <= 1e-5


def test_selected_schedule_end_to_end(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 13, 200000, 200000, (1.8, 20000))
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    s = sel.Selector()
    c, p, f = s.choose(m, n, dcp, dri)
    assert 0 <= c < sel.N_CLASSES and np.all(np.isfinite(f))
    run = sel.ScheduledSpMV(m, n, dcp, dri, dv, c)
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    run.set_x(DeviceArray.from_host(x))
    dy = DeviceArray(m, np.float32)
    run.step(dy)
    yo, S = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x)
    assert rel_err(dy.to_host().astype(np.float64), yo, S) <= 1e-5


def test_refined_choice_times_the_top_ranked(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 14, 100000, 100000, (1.9, 10000))
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    s = sel.Selector()
    ranked = s.rank_features(sel.features(m, n, dcp, dri))
    assert sorted(ranked) == list(range(sel.N_CLASSES))
    c1, _, _ = s.choose(m, n, dcp, dri)
    assert ranked[0] == c1
    c, times, f = s.choose_refined(m, n, dcp, dri, dv, top=3)
    # the HYB schedule joins the timed candidates when the 32768 longest columns hold at least
    # HYB_MIN_SHARE of the entries (they do on this Zipf matrix)
    share = sel.heavy_share(n, dcp)
    assert share >= sel.HYB_MIN_SHARE
    assert set(times) == set(ranked[:3]) | {sel.HYB_CLASS} and c == min(times, key=times.get)
    assert all(t > 0 for t in times.values())
    run = sel.ScheduledSpMV(m, n, dcp, dri, dv, c)
    x = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    run.set_x(DeviceArray.from_host(x))
    dy = DeviceArray(m, np.float32)
    run.step(dy)
    yo, S = ob.spmv_oracle(m, n, cp, ri, v.astype(np.float32), x)
    assert rel_err(dy.to_host().astype(np.float64), yo, S) <= 1e-5


def test_heavy_share_decides_the_hyb_candidate(gpu):
    """heavy_share is the share of entries in the 32768 longest columns; the rule adds the HYB
    schedule to the timed candidates only above HYB_MIN_SHARE (a 27-point stencil has no heavy columns: 32768 of 110592 equal
    columns hold 29.6 % of the entries)."""
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_STENCIL27, 15, 0, 0, (48,))
    dcp, dri, dv = DeviceArray.from_host(cp), DeviceArray.from_host(ri), DeviceArray.from_host(v)
    cnt = np.sort(np.diff(cp))[::-1]
    want = cnt[:32768].sum() / cp[n]
    assert abs(sel.heavy_share(n, dcp) - want) < 1e-12 and want < sel.HYB_MIN_SHARE
    s = sel.Selector()
    ranked = s.rank_features(sel.features(m, n, dcp, dri))
    _, times, _ = s.choose_refined(m, n, dcp, dri, dv, top=2)
    assert set(times) == set(ranked[:2])  # HYB only if the network ranks it in the top two


def test_native_selection_matches_features_and_network(gpu):
    s = sel.Selector()
    for case in range(5):
        m, n, cp, ri, v = list(feature_cases())[case]
        dcp, dri = DeviceArray.from_host(cp), DeviceArray.from_host(ri)
        c, p, f = s.choose(m, n, dcp, dri)
        cls, prob = s.mlp.predict(f[None, :])
        assert c == int(cls[0]) and abs(p - float(prob[0])) <= 1e-6
    # no network: the reference schedule, probability 0
    import ctypes as C
    out, pr = C.c_int(7), C.c_float(1.0)
    assert fl.lib.fl_select_schedule(None, m, n, len(ri), dcp.ptr, dri.ptr, C.byref(out), C.byref(pr), None) == 0
    assert out.value == 0 and pr.value == 0.0


def test_features_reject_malformed_csc(gpu):
    cp = np.array([0, 2, 3], np.int32)
    with pytest.raises(fl.FastLoadError):  # row index 9 outside a 5-row matrix
        sel.features(5, 2, DeviceArray.from_host(cp), DeviceArray.from_host(np.array([0, 9, 1], np.int32)))
    with pytest.raises(fl.FastLoadError):  # column pointers not monotone
        sel.features(5, 2, DeviceArray.from_host(np.array([0, 3, 2], np.int32)),
                     DeviceArray.from_host(np.array([0, 1], np.int32)))
