"""FL_VAL_PACKED on the GPU: one 32-bit word per entry holding the value (rounded to the mantissa
bits the slot leaves over) and the batched schedule's LDS slot. Pins: a one-entry-per-row matrix
times x = 1 returns every decoded value bit for bit against the oracle's rounding
(oracle/fastload_oracle.c or_pack_values); on larger matrices y matches the oracle of the decoded
values to the FP32 bar and the FP64 oracle within the format's stated bound, 2^-(m+1) of
sum_j |a_rj x_j| per row."""
import numpy as np
import pytest

import oracle_bind as ob
from gpu_util import rel_err

import fastload_amd as fl
from fastload_amd import DeviceArray

pytestmark = pytest.mark.gpu

SCHEDS = [("pb128", fl.KERNEL_PB, 128, 16384), ("pb256", fl.KERNEL_PB, 256, 4096), ("win128", fl.KERNEL_WIN, 128, 0),
          ("winr64", fl.KERNEL_WINR, 64, 0)]


def _run(M, x32, kernel):
    dx = DeviceArray.from_host(x32)
    dsx = DeviceArray(M.n, np.float32)
    M.permute_x(dx, dsx)
    dy = DeviceArray.from_host(np.full(max(M.m, 1), 5.0, np.float32))
    M.spmv(dsx, dy, kernel, fl.VAL_PACKED, assign=True)
    return dy.to_host()[:M.m].astype(np.float64)


@pytest.mark.parametrize("sched", SCHEDS, ids=[s[0] for s in SCHEDS])
def test_packed_values_decode_exactly(gpu, sched):
    name, kernel, bt, R = sched
    n = 200000
    rng = np.random.default_rng(3)
    rows = rng.permutation(n).astype(np.int32)
    # magnitudes over 24 binades: 5 exponent-code bits, so even PB with 256-tile batches (14 slot
    # bits) keeps 12 mantissa bits
    v = rng.choice([-1.0, 1.0], n) * (1.0 + rng.random(n)) * np.exp2(rng.integers(-12, 12, n))
    v[::97] = 0.0
    cp = np.arange(n + 1, dtype=np.int32)
    M = fl.FastLoadMatrix(n, n, cp, rows, v, width=64)
    M.prepare_schedule(kernel, bt, R)
    M.prepare(fl.VAL_PACKED)
    m_bits, e_bits, i_bits = M.pack_info(kernel)
    assert m_bits >= 12 and 1 + m_bits + e_bits + i_bits <= 32
    y = _run(M, np.ones(n, np.float32), kernel)
    want = np.zeros(n)
    want[rows] = ob.pack_values(v, m_bits)
    assert np.array_equal(y, want), name


@pytest.mark.parametrize("sched", SCHEDS, ids=[s[0] for s in SCHEDS])
@pytest.mark.parametrize("family", ["zipf", "stencil", "banded"])
def test_packed_accuracy_bound(gpu, sched, family):
    name, kernel, bt, R = sched
    m, n, cp, ri, v = {"zipf": lambda: fl.gen_matrix(fl.GEN_ZIPF, 21, 60000, 90000, (1.8, 3000)),
                       "stencil": lambda: fl.gen_matrix(fl.GEN_STENCIL27, 21, 0, 0, (24,)),
                       "banded": lambda: fl.gen_matrix(fl.GEN_BANDED, 21, 50000, 50000, (40,))}[family]()
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare(fl.VAL_PACKED)  # before the schedule: packed when the schedule is prepared
    M.prepare_schedule(kernel, bt, R)
    m_bits = M.pack_info(kernel)[0]
    x = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    y = _run(M, x, kernel)
    y_dec, s_dec = ob.spmv_oracle(m, n, cp, ri, ob.pack_values(v, m_bits), x)
    assert rel_err(y, y_dec, s_dec) <= 1e-5
    y64, s64 = ob.spmv_oracle(m, n, cp, ri, v, x.astype(np.float64))
    bound = 2.0 ** -(m_bits + 1) * (1 + 1e-6) + 1e-6
    assert rel_err(y, y64, s64) <= bound, (name, family, m_bits)
    err_l2 = np.linalg.norm(y - y64) / np.linalg.norm(y64)
    print(f"{family} {name}: m = {m_bits} bits, rel L2 error {err_l2:.2e}", flush=True)
    assert err_l2 <= 1e-4


def test_packed_refusals(gpu):
    m, n, cp, ri, v = fl.gen_matrix(fl.GEN_ZIPF, 22, 20000, 20000, (1.8, 1000))
    M = fl.FastLoadMatrix(m, n, cp, ri, v, width=64)
    M.prepare_schedule(fl.KERNEL_PB, 128, 4096)
    dsx = DeviceArray.from_host(np.ones(n, np.float32))
    dy = DeviceArray(m, np.float32)
    with pytest.raises(fl.FastLoadError):  # not prepared
        M.spmv(dsx, dy, fl.KERNEL_PB, fl.VAL_PACKED, assign=True)
    M.prepare(fl.VAL_PACKED)
    for kernel in (fl.KERNEL_DROPIN, fl.KERNEL_FAST):  # not a batched schedule
        with pytest.raises(fl.FastLoadError):
            M.spmv(dsx, dy, kernel, fl.VAL_PACKED, assign=True)
    # a value range of 2^200 needs 8 exponent bits: with PB's 13 slot bits only 10 mantissa bits
    # would remain, so PB keeps no packed words
    w = v * np.exp2(np.random.default_rng(5).integers(-100, 100, v.size))
    W = fl.FastLoadMatrix(m, n, cp, ri, w, width=64)
    W.prepare_schedule(fl.KERNEL_PB, 128, 4096)
    W.prepare(fl.VAL_PACKED)
    with pytest.raises(fl.FastLoadError):
        W.pack_info(fl.KERNEL_PB)
    with pytest.raises(fl.FastLoadError):
        W.spmv(dsx, dy, fl.KERNEL_PB, fl.VAL_PACKED, assign=True)
