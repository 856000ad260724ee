/*
 * fastload_gen.h -- seeded synthetic CSC matrices and vectors for tests and benchmarks.
 *
 * The reference ships one matrix (Data/1138_bus.mtx) and ad-hoc random generators inside its
 * drivers (test.cu:30-43, benchmarks.cu:8-21); these generators produce the workloads the
 * benchmark configurations describe. Output is CSC with rows ascending inside each column,
 * no duplicate (row, column) pairs; arrays are malloc-allocated (free with fl_free).
 * Every column draws from its own counter-based random stream, so results do not depend on
 * the number of host threads.
 */
#ifndef FASTLOAD_GEN_H
#define FASTLOAD_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Families and their params[]:
 *  FL_GEN_SYMGEOM   structurally symmetric n x n, full diagonal; off-diagonal degree per column
 *                   geometric on {1,2,..} with mean p0 (default 2.6), capped at p1 (default 17);
 *                   values U[-1,1] (symmetric pairs share a value), diagonal U[-1,1]+4.
 *  FL_GEN_ZIPF      column counts Zipf: P(k) ~ k^-p0 on 1..min(p1, m); rows uniform without
 *                   repeats; values N(0,1).
 *  FL_GEN_STENCIL27 27-point stencil on a p0^3 grid (m = n = p0^3); |v| ~ U[0.5,1.5], random
 *                   sign.
 *  FL_GEN_BANDED    column j holds every row in [j - p0/2, j + p0/2] inside the matrix;
 *                   values N(0,1).
 *  FL_GEN_BLOCKDIAG diagonal blocks of size p0; each column holds min(p1, block) random rows of
 *                   its block (p1 <= 0: the whole block); values N(0,1).
 *  FL_GEN_UNIFORM   each column holds Binomial(m, p0) random rows (Poisson approximation);
 *                   values N(0,1).
 *  FL_GEN_ZIPF_DENSE Zipf (p0, p1) plus a fraction p2 of columns holding min(p3, m) rows.
 */
#define FL_GEN_SYMGEOM 0
#define FL_GEN_ZIPF 1
#define FL_GEN_STENCIL27 2
#define FL_GEN_BANDED 3
#define FL_GEN_BLOCKDIAG 4
#define FL_GEN_UNIFORM 5
#define FL_GEN_ZIPF_DENSE 6

int fl_gen_matrix(int family, uint64_t seed, int m, int n, const double *params, int nparams, int **colptr,
                  int **rowidx, double **val, int *nnz);

/* Vectors: FL_VEC_NORMAL N(0,1); FL_VEC_UNIFORM U[-1,1]; FL_VEC_MOD10 out[i] = i % 10
 * (the integer data of src/main.cu:93-96). */
#define FL_VEC_NORMAL 0
#define FL_VEC_UNIFORM 1
#define FL_VEC_MOD10 2
int fl_gen_vector(int dist, uint64_t seed, int count, double *out);

#ifdef __cplusplus
}
#endif

#endif /* FASTLOAD_GEN_H */
