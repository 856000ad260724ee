"""Selector oracle -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the learned schedule selector's two computations, used exclusively as the
checker of the HIP kernels (fl_features, fl_mlp_*): only tests/ import it; the fastload_amd
package never does. There is no reference counterpart (the reference runs one fixed schedule; its
unused column statistics are src/auxiliary.h:6-25).

  features()      the 40 features of a CSC matrix, as documented for fl_features
  mlp_forward()   the 3-layer perceptron forward pass and mean soft-target cross-entropy
  mlp_backward()  its gradients, in FP64
  bf16()          round-to-nearest-even to bfloat16 (returned as float32 values)

mlp_forward/mlp_backward take `rounding`: with True, operands are rounded to BF16 at the points
where the HIP kernels round them (inputs and weights of every matrix product, the hidden
activations, the output and hidden-layer error signals entering a product); with False the
computation is exact FP64. Pins (tests/test_host.py): features against per-column/per-row Python
loops on small matrices and the closed form of a band matrix; the FP64 backward against central
finite differences of the FP64 forward; bf16() against the bit patterns of hand-chosen values
(ties to even, carries into the exponent).
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even); NaN stays NaN."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    nan = np.isnan(a)
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def features(m, n, colptr, rowidx):
    """fl_features restated: 32-bin histogram of column lengths (bin 0 empty, bin b >= 1 holds
    lengths in [2^(b-1), 2^b), bin 31 open) as fractions of n, then log2 m/n/nnz and row
    statistics scaled as documented in include/fastload.h."""
    colptr = np.asarray(colptr, np.int64)
    rowidx = np.asarray(rowidx, np.int64)
    nnz = int(colptr[n]) if n else 0
    f = np.zeros(40, np.float64)
    lens = np.diff(colptr[:n + 1])
    if n:
        bins = np.where(lens > 0, np.minimum(31, np.frexp(np.maximum(lens, 1).astype(np.float64))[1]), 0)
        f[:32] = np.bincount(bins, minlength=32)[:32] / n
    rc = np.bincount(rowidx[:nnz], minlength=max(m, 1))[:m] if m else np.zeros(0)
    mean = nnz / m if m else 0.0
    var = max(float((rc.astype(np.float64) ** 2).mean()) - mean * mean, 0.0) if m else 0.0
    f[32] = np.log2(m + 1.0) / 24.0
    f[33] = np.log2(n + 1.0) / 24.0
    f[34] = np.log2(nnz + 1.0) / 32.0
    f[35] = np.log2(mean + 1.0) / 8.0
    f[36] = np.sqrt(var) / mean if mean > 0 else 0.0
    f[37] = np.log2(rc.max() / mean + 1.0) / 8.0 if mean > 0 else 0.0
    f[38] = float((rc == 0).mean()) if m else 0.0
    ne = lens > 0
    if ne.any():
        first = rowidx[colptr[:n][ne]]
        last = rowidx[colptr[1:n + 1][ne] - 1]
        f[39] = float(((last - first) / max(m, 1)).mean())
    return f.astype(np.float32)


def _r(a, rounding):
    return bf16(a).astype(np.float64) if rounding else np.asarray(a, np.float64)


def unpack(params, dims):
    out, o = [], 0
    for l in range(3):
        K, N = dims[l], dims[l + 1]
        out.append(np.asarray(params[o:o + K * N], np.float64).reshape(K, N))
        o += K * N
        out.append(np.asarray(params[o:o + N], np.float64))
        o += N
    return out


def mlp_forward(params, dims, X, T=None, rounding=False):
    """Returns (logits, cache, loss); loss is None when T is None."""
    W1, b1, W2, b2, W3, b3 = unpack(params, dims)
    Xr = _r(X, rounding)
    Wr = [_r(W, rounding) for W in (W1, W2, W3)]
    Z1 = Xr @ Wr[0] + b1
    H1 = _r(np.maximum(Z1, 0.0), rounding)
    Z2 = H1 @ Wr[1] + b2
    H2 = _r(np.maximum(Z2, 0.0), rounding)
    Z3 = H2 @ Wr[2] + b3
    cache = dict(Xr=Xr, Wr=Wr, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3)
    loss = None
    if T is not None:
        T = np.asarray(T, np.float64)
        zs = Z3 - Z3.max(axis=1, keepdims=True)
        lp = zs - np.log(np.exp(zs).sum(axis=1, keepdims=True))
        loss = float(-(T * lp).sum(axis=1).mean())
        cache["P"] = np.exp(lp)
    return Z3, cache, loss


def mlp_backward(params, dims, X, T, rounding=False):
    """Gradients of the mean soft-target cross-entropy, flattened in parameter order
    (W1, b1, W2, b2, W3, b3); returns (grads, loss)."""
    _, c, loss = mlp_forward(params, dims, X, T, rounding)
    B = X.shape[0]
    T = np.asarray(T, np.float64)
    dZ3 = (c["P"] - T) / B
    dZ3r = _r(dZ3, rounding)
    gW3 = c["H2"].T @ dZ3r
    gb3 = dZ3.sum(axis=0)
    dH2 = dZ3r @ c["Wr"][2].T
    dZ2 = dH2 * (c["Z2"] > 0)
    dZ2r = _r(dZ2, rounding)
    gW2 = c["H1"].T @ dZ2r
    gb2 = dZ2.sum(axis=0)
    dH1 = dZ2r @ c["Wr"][1].T
    dZ1 = dH1 * (c["Z1"] > 0)
    dZ1r = _r(dZ1, rounding)
    gW1 = c["Xr"].T @ dZ1r
    gb1 = dZ1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return g, loss
