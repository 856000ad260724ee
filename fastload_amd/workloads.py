"""Seeded synthetic workloads the benchmark and the tests share (host CSC arrays from the
library's own generators, include/fastload_gen.h):

  symgeom(seed)         1138 x 1138 structurally symmetric matrix with the statistics of the
                        reference's only input, Data/1138_bus.mtx (full diagonal, off-diagonal
                        column counts geometric with mean 2.6 capped at 17, values U[-1,1] with
                        the diagonal shifted by +4); x U[-1,1]
  powerlaw(n, seed)     n x n, column counts Zipf(1.8423) truncated at 50,000 (mean 16), rows
                        uniform without repeats per column, values N(0,1); x N(0,1)
  stencil(g)            27-point stencil on a g^3 grid, |v| U[0.5,1.5] with random sign
  laplacian27(g)        the same 27-point structure with 26 on the diagonal and -1 elsewhere:
                        27 I - T(x)T(x)T with T = tridiag(1, 1, 1), whose eigenvalues lie in
                        (0, 54), so the matrix is symmetric positive definite (a CG system)
  batch(count, seed)    `count` square matrices, n log-uniform in [1e3, 2e5], column counts
                        Zipf(2.0) truncated at n (mean ~7.8), values N(0,1)
"""
from __future__ import annotations

import numpy as np

from .api import GEN_STENCIL27, GEN_SYMGEOM, GEN_ZIPF, VEC_NORMAL, VEC_UNIFORM, gen_matrix, gen_vector

POWERLAW_ZIPF = 1.8423  # truncated at 50,000 this exponent has mean 16.0 (1.8 would give 20.8)
POWERLAW_MAX = 50000


def symgeom(seed: int = 1138):
    m, n, cp, ri, v = gen_matrix(GEN_SYMGEOM, seed, 1138, 1138)
    return m, n, cp, ri, v, gen_vector(VEC_UNIFORM, seed + 1, n)


def powerlaw(n: int = 4194304, seed: int = 42):
    m, n, cp, ri, v = gen_matrix(GEN_ZIPF, seed, n, n, (POWERLAW_ZIPF, POWERLAW_MAX))
    return m, n, cp, ri, v, gen_vector(VEC_NORMAL, seed + 1, n)


def stencil(g: int = 128, seed: int = 128):
    m, n, cp, ri, v = gen_matrix(GEN_STENCIL27, seed, 0, 0, (g,))
    return m, n, cp, ri, v, gen_vector(VEC_NORMAL, seed + 1, n)


def laplacian27(g: int = 128, seed: int = 128):
    """(m, n, colptr, rowidx, val, b): the symmetric positive definite 27-point matrix on g^3 and
    a right-hand side N(0,1)."""
    m, n, cp, ri, _ = gen_matrix(GEN_STENCIL27, seed, 0, 0, (g,))
    col = np.repeat(np.arange(n, dtype=np.int32), np.diff(cp))
    v = np.where(ri == col, 26.0, -1.0)
    return m, n, cp, ri, v, gen_vector(VEC_NORMAL, seed + 1, n)


def batch_sizes(count: int = 512, seed: int = 512) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(1e3), np.log(2e5), count)).astype(int)


def batch(count: int = 512, seed: int = 512):
    """[(m, n, colptr, rowidx, val), ...] of the batched configuration."""
    return [gen_matrix(GEN_ZIPF, seed * 100000 + i, int(n), int(n), (2.0, int(n)))
            for i, n in enumerate(batch_sizes(count, seed))]
