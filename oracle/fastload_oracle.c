/*
 * FastLoad CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference FastLoad algorithm (skyler-ruiter/FastLoad), used
 * exclusively as the checker for the HIP product path: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library. Nothing in the fastload_amd package
 * links, imports or calls it, and the product path never falls back to it.
 *
 * Every function cites the reference file:line it restates. Where the reference's GPU path
 * and its CPU twins (src/coo2csc.h) differ, the GPU path is canonical:
 *   - the column sort is STABLE (thrust::sort_by_key on int keys, src/ColSort_GPU.h:101),
 *     not the unstable std::sort of src/coo2csc.h:100;
 *   - the tile width used for level/tile enumeration and classification is 32
 *     (src/formatTransform_GPU.h:94,119,133,273,277); the oracle takes the width as a
 *     parameter W so the same code also describes the 64-wide schedule variant.
 * Reference defects are not reproduced (uninitialised d_segOff, missing bounds checks,
 * quadratic prefix loops): the oracle computes the intended values.
 *
 * Pins (tests/test_oracle.py): Data/1138_bus.mtx facts and scipy.io.mmread for the loader;
 * scipy.sparse tocsc() for csr2csc; numpy argsort(kind="stable") for the column sort; the
 * compiled reference CPU twins (oracle/_ref, built from the reference's own headers) for
 * formattransation / tile_classification / FastLoad_cpu, bit-exact; closed-form tile counts
 * and the reference's own invariants (tile_ptr[tilenum]==nnz, sum of runs == nnz); integer
 * data (values i%10, x i%10) for exact y; the 9x9 hand matrix of spmv.cu:140-150; and
 * math.fsum for the compensated FP64 oracle.
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_ERR_OPEN (-1)      /* AUX:81-82 */
#define OR_ERR_BANNER (-2)    /* AUX:84-88 */
#define OR_ERR_TRUNC (-3)     /* entry lines missing (reference ignores fscanf result, AUX:126) */
#define OR_ERR_SIZE (-4)      /* AUX:96-98 */
#define OR_ERR_RANGE (-5)     /* index outside [1,m]x[1,n] (reference: out-of-bounds write) */
#define OR_ERR_UNSUPPORTED (-6) /* dense "array" files: the reference loader cannot read them */
#define OR_ERR_ALLOC (-7)
#define OR_ERR_INVALID (-8)

void or_free(void *p) { free(p); }

/* ------------------------------------------------------------------------------------------
 * Matrix Market banner + size line: mmio.h:172-255 (mm_read_banner), mmio.h:142-170
 * (mm_read_mtx_crd_size). Tokens after the banner are case-insensitive.
 * type: 'R' real, 'C' complex, 'P' pattern, 'I' integer; sym: 'G','S','H','K'.
 * ---------------------------------------------------------------------------------------- */
static void lower(char *s) { for (; *s; ++s) *s = (char)tolower((unsigned char)*s); }

static int read_banner(FILE *f, char *type, char *sym, int *dense)
{
    char line[1025], banner[64], mtx[64], crd[64], dt[64], ss[64];
    if (!fgets(line, sizeof line, f)) return OR_ERR_BANNER;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dt, ss) != 5) return OR_ERR_BANNER;
    lower(mtx); lower(crd); lower(dt); lower(ss);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return OR_ERR_BANNER;
    if (strcmp(mtx, "matrix") != 0) return OR_ERR_BANNER;
    if (strcmp(crd, "coordinate") == 0) *dense = 0;
    else if (strcmp(crd, "array") == 0) *dense = 1;
    else return OR_ERR_BANNER;
    if (!strcmp(dt, "real")) *type = 'R';
    else if (!strcmp(dt, "complex")) *type = 'C';
    else if (!strcmp(dt, "pattern")) *type = 'P';
    else if (!strcmp(dt, "integer")) *type = 'I';
    else return OR_ERR_BANNER;
    if (!strcmp(ss, "general")) *sym = 'G';
    else if (!strcmp(ss, "symmetric")) *sym = 'S';
    else if (!strcmp(ss, "hermitian")) *sym = 'H';
    else if (!strcmp(ss, "skew-symmetric")) *sym = 'K';
    else return OR_ERR_BANNER;
    return OR_OK;
}

static int read_size(FILE *f, int *m, int *n, int *nz)
{
    char line[1025];
    *m = *n = *nz = 0;
    do {
        if (!fgets(line, sizeof line, f)) return OR_ERR_SIZE;
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", m, n, nz) == 3) return OR_OK;
    for (;;) {
        int got = fscanf(f, "%d %d %d", m, n, nz);
        if (got == EOF) return OR_ERR_SIZE;
        if (got == 3) return OR_OK;
        return OR_ERR_SIZE; /* a non-numeric token would loop forever in the reference */
    }
}

/*
 * mmio_allinone (AUX:65-232): .mtx -> 0-based CSR.
 *   - pattern -> 1.0, integer -> (double)int, complex -> real part (AUX:128-146)
 *   - symmetric/hermitian: every off-diagonal entry (r,c) is mirrored (AUX:100-104,160-204);
 *     skew-symmetric is treated as general (no mirror), as in the reference.
 *   - CSR order: rows by index; inside a row, entries in the order they are inserted while
 *     scanning the file, the mirror (c,r) inserted into row c when its source line is read
 *     (AUX:180-203). Columns inside a row are NOT sorted.
 * Returns 0 or a negative error code; on success the three CSR arrays are malloc-allocated.
 */
int or_mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric, int **rowptr, int **colidx,
                     double **val, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return OR_ERR_OPEN;
    char type = 0, sym = 0;
    int dense = 0;
    int rc = read_banner(f, &type, &sym, &dense);
    if (rc) { fclose(f); return rc; }
    if (dense) { fclose(f); return OR_ERR_UNSUPPORTED; }
    int mm, nn, nzr;
    rc = read_size(f, &mm, &nn, &nzr);
    if (rc) { fclose(f); return rc; }
    if (mm < 0 || nn < 0 || nzr < 0) { fclose(f); return OR_ERR_SIZE; }
    int is_sym = (sym == 'S' || sym == 'H');

    int *ri = (int *)malloc(sizeof(int) * (size_t)(nzr ? nzr : 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(nzr ? nzr : 1));
    double *vv = (double *)malloc(sizeof(double) * (size_t)(nzr ? nzr : 1));
    long long *cnt = (long long *)calloc((size_t)mm + 1, sizeof(long long));
    if (!ri || !ci || !vv || !cnt) { fclose(f); free(ri); free(ci); free(vv); free(cnt); return OR_ERR_ALLOC; }

    for (int i = 0; i < nzr; ++i) {
        int a, b, iv, got;
        double fv = 0.0, fim;
        if (type == 'R') { got = fscanf(f, "%d %d %lg", &a, &b, &fv); rc = (got == 3); }
        else if (type == 'C') { got = fscanf(f, "%d %d %lg %lg", &a, &b, &fv, &fim); rc = (got == 4); }
        else if (type == 'I') { got = fscanf(f, "%d %d %d", &a, &b, &iv); fv = iv; rc = (got == 3); }
        else { got = fscanf(f, "%d %d", &a, &b); fv = 1.0; rc = (got == 2); }
        if (!rc) { rc = OR_ERR_TRUNC; goto fail; }
        if (a < 1 || a > mm || b < 1 || b > nn) { rc = OR_ERR_RANGE; goto fail; }
        ri[i] = a - 1;
        ci[i] = b - 1;
        vv[i] = fv;
        cnt[a - 1]++;
    }
    fclose(f);
    f = NULL;
    if (is_sym) {
        for (int i = 0; i < nzr; ++i) {
            if (ri[i] != ci[i]) {
                if (ci[i] >= mm || ri[i] >= nn) { rc = OR_ERR_RANGE; goto fail; } /* mirror needs a square index space */
                cnt[ci[i]]++;
            }
        }
    }
    {
        long long total = 0;
        for (int r = 0; r <= mm; ++r) { long long c = cnt[r]; cnt[r] = total; total += c; }
        if (total > 0x7fffffffLL) { rc = OR_ERR_SIZE; goto fail; }
        int nz = (int)total;
        int *rp = (int *)malloc(sizeof(int) * ((size_t)mm + 1));
        int *cc = (int *)malloc(sizeof(int) * (size_t)(nz ? nz : 1));
        double *vo = (double *)malloc(sizeof(double) * (size_t)(nz ? nz : 1));
        long long *fill = (long long *)calloc((size_t)mm + 1, sizeof(long long));
        if (!rp || !cc || !vo || !fill) { free(rp); free(cc); free(vo); free(fill); rc = OR_ERR_ALLOC; goto fail; }
        for (int r = 0; r <= mm; ++r) rp[r] = (int)cnt[r];
        for (int i = 0; i < nzr; ++i) {
            int r = ri[i], c = ci[i];
            long long off = cnt[r] + fill[r]++;
            cc[off] = c;
            vo[off] = vv[i];
            if (is_sym && r != c) {
                off = cnt[c] + fill[c]++;
                cc[off] = r;
                vo[off] = vv[i];
            }
        }
        free(fill);
        *m = mm; *n = nn; *nnz = nz; *isSymmetric = is_sym;
        *rowptr = rp; *colidx = cc; *val = vo;
    }
    free(ri); free(ci); free(vv); free(cnt);
    return OR_OK;
fail:
    if (f) fclose(f);
    free(ri); free(ci); free(vv); free(cnt);
    return rc;
}

/* exclusive_scan (AUX:29-44) over length entries, in place. */
void or_exclusive_scan(int *a, int length)
{
    int run = 0;
    for (int i = 0; i < length; ++i) { int v = a[i]; a[i] = run; run += v; }
}

/* ------------------------------------------------------------------------------------------
 * csr2csc (COO:11-72; GPU: cusparseCsr2cscEx2 + nnzpercolSegmentReduce, C2C:24-32,88-109).
 * Counting transpose: rows visited ascending, entries of a row in CSR order, so inside each
 * column entries appear by ascending row (ties by CSR position).
 * ---------------------------------------------------------------------------------------- */
void or_csr2csc(int m, int n, int nnz, const int *rowptr, const int *colidx, const double *val,
                int *colptr, int *rowidx, double *cval, int *nnzpercol)
{
    (void)nnz;
    memset(colptr, 0, sizeof(int) * ((size_t)n + 1));
    for (int r = 0; r < m; ++r)
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) colptr[colidx[j] + 1]++;
    for (int c = 0; c < n; ++c) { nnzpercol[c] = colptr[c + 1]; colptr[c + 1] += colptr[c]; }
    int *next = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
    memcpy(next, colptr, sizeof(int) * (size_t)n);
    for (int r = 0; r < m; ++r)
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            int dst = next[colidx[j]]++;
            rowidx[dst] = r;
            cval[dst] = val[j];
        }
    free(next);
}

/* ------------------------------------------------------------------------------------------
 * ColSort (CS:46-137): perm = stable ascending sort of column lengths (thrust::sort_by_key,
 * CS:93-101); sortnnz[i] = c[perm[i]]; sortx[i] = x[perm[i]] (Pre4sortedx, CS:9-19);
 * the columns' (row, val) lists concatenated in perm order, each keeping its in-column order
 * (Pre4SortIdxVal, CS:21-43). The stable sort is a counting sort over the key range.
 * ---------------------------------------------------------------------------------------- */
void or_colsort(int n, const int *colptr, const int *rowidx, const double *val, const double *x,
                int *perm, int *sortnnz, int *tmp_row, double *tmp_val, double *sortx)
{
    int maxk = 0;
    for (int c = 0; c < n; ++c) { int k = colptr[c + 1] - colptr[c]; if (k > maxk) maxk = k; }
    int *bucket = (int *)calloc((size_t)maxk + 2, sizeof(int));
    for (int c = 0; c < n; ++c) bucket[colptr[c + 1] - colptr[c] + 1]++;
    for (int k = 0; k <= maxk; ++k) bucket[k + 1] += bucket[k];
    for (int c = 0; c < n; ++c) perm[bucket[colptr[c + 1] - colptr[c]]++] = c; /* stable */
    free(bucket);
    long long pos = 0;
    for (int i = 0; i < n; ++i) {
        int c = perm[i];
        int k = colptr[c + 1] - colptr[c];
        sortnnz[i] = k;
        if (sortx) sortx[i] = x[c];
        for (int d = 0; d < k; ++d, ++pos) {
            tmp_row[pos] = rowidx[colptr[c] + d];
            if (tmp_val) tmp_val[pos] = val[colptr[c] + d];
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Levels (SegmentReduce FT:8-17, prepare FT:21-57, countNonZeroElements FT:60-71;
 * CPU twin COO:142-186): r[0]=s[0], r[i]=s[i]-s[i-1]; every i with r[i]!=0 is a level
 * (col=i, base=s[i-1] or 0, tall=r[i]).
 * Tiles (countTileNum FT:85-98, auxiliaryTrans FT:113-150; COO:197-250): per level, per
 * chunk j of W consecutive sorted columns (last chunk short), per depth e < tall: one tile.
 * ---------------------------------------------------------------------------------------- */
int or_count_levels_tiles(int n, const int *sortnnz, int W, int *L, long long *tilenum)
{
    if (W <= 0) return OR_ERR_INVALID;
    int nl = 0;
    long long tn = 0;
    for (int i = 0; i < n; ++i) {
        int r = (i == 0) ? sortnnz[0] : sortnnz[i] - sortnnz[i - 1];
        if (r < 0) return OR_ERR_INVALID; /* lengths must be sorted ascending */
        if (r != 0) {
            long long chunks = ((long long)(n - i) + W - 1) / W;
            tn += chunks * r;
            nl++;
        }
    }
    *L = nl;
    *tilenum = tn;
    return OR_OK;
}

/*
 * formatTransform data layout (auxiliaryTrans FT:113-150, scan #2 FT:435-439, dataTrans
 * FT:152-171, PtrReduce FT:173-185; CPU twin formattransation COO:132-312):
 *   tile_ptr = exclusive scan of tile_len (tile_ptr[tilenum] == nnz);
 *   sortrowidx[tile_ptr[t]+l] = tmp_row[P[tile_colidx[t]+l] + depth(t)], P = exclusive scan
 *   of sortnnz (scan #1, FT:399-406).
 * Outputs sized: lvl_* [L], tile_ptr [tilenum+1], tile_len/colidx/depth [tilenum], data [nnz].
 */
int or_format_transform(int n, int nnz, const int *sortnnz, const int *tmp_row, const double *tmp_val,
                        int W, int *lvl_col, int *lvl_base, int *lvl_tall, int *tile_ptr, int *tile_len,
                        int *tile_colidx, int *tile_depth, int *sortrowidx, double *sortval)
{
    if (W <= 0) return OR_ERR_INVALID;
    long long *P = (long long *)malloc(sizeof(long long) * ((size_t)n + 1));
    if (!P) return OR_ERR_ALLOC;
    P[0] = 0;
    for (int i = 0; i < n; ++i) P[i + 1] = P[i] + sortnnz[i];
    if (P[n] != nnz) { free(P); return OR_ERR_INVALID; }

    int k = 0;
    for (int i = 0; i < n; ++i) {
        int r = (i == 0) ? sortnnz[0] : sortnnz[i] - sortnnz[i - 1];
        if (r != 0) {
            lvl_col[k] = i;
            lvl_base[k] = (i == 0) ? 0 : sortnnz[i - 1];
            lvl_tall[k] = r;
            k++;
        }
    }
    long long t = 0;
    for (int q = 0; q < k; ++q) {
        int col = lvl_col[q], width = n - col;
        int chunks = (width + W - 1) / W;
        for (int j = 0; j < chunks; ++j) {
            int len = (j == chunks - 1) ? width - (chunks - 1) * W : W;
            for (int e = 0; e < lvl_tall[q]; ++e, ++t) {
                tile_len[t] = len;
                tile_colidx[t] = col + W * j;
                tile_depth[t] = lvl_base[q] + e;
            }
        }
    }
    long long run = 0;
    for (long long u = 0; u < t; ++u) { tile_ptr[u] = (int)run; run += tile_len[u]; }
    tile_ptr[t] = (int)run;
    if (run != nnz) { free(P); return OR_ERR_INVALID; }
    for (long long u = 0; u < t; ++u) {
        for (int l = 0; l < tile_len[u]; ++l) {
            long long src = P[tile_colidx[u] + l] + tile_depth[u];
            sortrowidx[tile_ptr[u] + l] = tmp_row[src];
            if (sortval && tmp_val) sortval[tile_ptr[u] + l] = tmp_val[src];
        }
    }
    free(P);
    return OR_OK;
}

/*
 * Segments and tile format (Pre4CountRow FT:188-208, scans #3/#4, Pre4segOff FT:210-234,
 * Pre4addone FT:235-244, Pre4SegIndex FT:246-261, Pre4tileclass FT:263-285; CPU twin
 * tile_classification COO:314-448):
 *   segments = maximal runs of equal sortrowidx inside a tile;
 *   sortrowindex[g] = run length if g heads a run, else 0;
 *   countrow = exclusive scan of per-tile run counts (tilenum+1), segsum = total runs;
 *   segmentoffset[s] = length of run s (segsum+1 entries, last 0);
 *   tile_format = 1 if len==W and one run, 2 if len==W and several runs, else 0.
 */
int or_classify(int tilenum, int nnz, int W, const int *tile_ptr, const int *tile_len, const int *sortrowidx,
                int *countrow, int *segmentoffset, int *sortrowindex, int *tile_format, int *segsum)
{
    memset(sortrowindex, 0, sizeof(int) * (size_t)(nnz > 0 ? nnz : 0));
    int total = 0;
    for (int t = 0; t < tilenum; ++t) {
        int start = tile_ptr[t], len = tile_len[t];
        int runs = 0;
        int head = start;
        for (int g = start; g < start + len; ++g) {
            int is_head = (g == start) || (sortrowidx[g] != sortrowidx[g - 1]);
            if (is_head) {
                if (g != start) {
                    sortrowindex[head] = g - head;
                    if (segmentoffset) segmentoffset[total + runs - 1] = g - head;
                }
                head = g;
                runs++;
            }
        }
        if (len > 0) {
            sortrowindex[head] = start + len - head;
            if (segmentoffset) segmentoffset[total + runs - 1] = start + len - head;
        }
        countrow[t] = total;
        tile_format[t] = (len == W) ? (runs == 1 ? 1 : 2) : 0;
        total += runs;
    }
    countrow[tilenum] = total;
    if (segmentoffset) segmentoffset[total] = 0;
    *segsum = total;
    return OR_OK;
}

/* ------------------------------------------------------------------------------------------
 * FastLoad_cpu (FC:19-31): serial loop over tiles in tile order,
 *   y[sortrowidx[j]] += sortval[j] * x[tile_colidx[t] + (j - start) % len]   (x = sortx).
 * y is accumulated in place (the caller zeroes it). The exact-equality check of FC:33-45 is a
 * separate concern (the tests compare results themselves).
 * ---------------------------------------------------------------------------------------- */
void or_fastload_cpu_f64(int tilenum, const int *tile_ptr, const int *tile_len, const int *tile_colidx,
                         const int *sortrowidx, const double *sortval, const double *sortx, double *y)
{
    for (int t = 0; t < tilenum; ++t) {
        int start = tile_ptr[t], stop = tile_ptr[t + 1], len = tile_len[t];
        for (int j = start; j < stop; ++j) y[sortrowidx[j]] += sortval[j] * sortx[tile_colidx[t] + (j - start) % len];
    }
}

void or_fastload_cpu_f32(int tilenum, const int *tile_ptr, const int *tile_len, const int *tile_colidx,
                         const int *sortrowidx, const float *sortval, const float *sortx, float *y)
{
    for (int t = 0; t < tilenum; ++t) {
        int start = tile_ptr[t], stop = tile_ptr[t + 1], len = tile_len[t];
        for (int j = start; j < stop; ++j) y[sortrowidx[j]] += sortval[j] * sortx[tile_colidx[t] + (j - start) % len];
    }
}

/* Golden y of MAIN:98-108 (CSR rows) and MAIN:110-120 (CSC columns), plain FP64. */
void or_spmv_csr_f64(int m, const int *rowptr, const int *colidx, const double *val, const double *x, double *y)
{
    for (int r = 0; r < m; ++r) {
        double s = 0.0;
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) s += val[j] * x[colidx[j]];
        y[r] = s;
    }
}

void or_spmv_csc_f64(int m, int n, const int *colptr, const int *rowidx, const double *val, const double *x, double *y)
{
    memset(y, 0, sizeof(double) * (size_t)m);
    for (int c = 0; c < n; ++c)
        for (int j = colptr[c]; j < colptr[c + 1]; ++j) y[rowidx[j]] += val[j] * x[c];
}

/*
 * FP64 oracle with compensated accumulation (Ogita, Rump, Oishi, "Accurate sum and dot
 * product", SIAM J. Sci. Comput. 26(6), 2005: Dot2 = TwoProduct via FMA + TwoSum), per row
 * over the CSC entries. The result is as accurate as if computed in twice the working
 * precision and then rounded. scale[r] = sum_j |a_rj x_j| is the condition-aware scale used
 * for relative tolerances. Values may be given in float (val32 != NULL) to evaluate the
 * exact product of the stored FP32 data.
 */
void or_spmv_oracle(int m, int n, const int *colptr, const int *rowidx, const double *val64, const float *val32,
                    const double *x64, const float *x32, double *y, double *scale)
{
    double *s = (double *)calloc((size_t)m, sizeof(double));
    double *c = (double *)calloc((size_t)m, sizeof(double));
    double *a = (double *)calloc((size_t)m, sizeof(double));
    for (int col = 0; col < n; ++col) {
        double xv = x64 ? x64[col] : (double)x32[col];
        for (int j = colptr[col]; j < colptr[col + 1]; ++j) {
            int r = rowidx[j];
            double v = val64 ? val64[j] : (double)val32[j];
            double p = v * xv;
            double pe = fma(v, xv, -p);
            double t = s[r] + p;
            double z = t - s[r];
            double se = (s[r] - (t - z)) + (p - z);
            s[r] = t;
            c[r] += se + pe;
            a[r] += fabs(p);
        }
    }
    for (int r = 0; r < m; ++r) {
        y[r] = s[r] + c[r];
        if (scale) scale[r] = a[r];
    }
    free(s); free(c); free(a);
}

/* ------------------------------------------------------------------------------------------
 * BF16 helpers for the precision variants (no reference counterpart: the reference is FP64
 * only). bf16 = the upper 16 bits of an FP32 after round-to-nearest-even; NaN kept quiet.
 * hi = bf16(v), lo = bf16(v - hi) with v the FP32 value (v - hi is exact in FP32).
 * ---------------------------------------------------------------------------------------- */
static uint16_t bf16_rne(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

static float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

void or_bf16_split(int count, const float *v, uint16_t *hi, uint16_t *lo)
{
    for (int i = 0; i < count; ++i) {
        uint16_t h = bf16_rne(v[i]);
        hi[i] = h;
        if (lo) lo[i] = bf16_rne(v[i] - bf16_to_f32(h));
    }
}

/* ------------------------------------------------------------------------------------------
 * Packed value storage of the batched schedules (no reference counterpart): each value rounded
 * to nearest even from FP64 to m explicit mantissa bits (1 <= m <= 23), magnitudes below the FP32
 * normal range (exponent < -126) stored as zero, the sign of zero dropped. Returns the value the
 * kernels decode, as a float (exact: m <= 23 and the exponent stays inside the FP32 range for the
 * values the format accepts, exponent < 127).
 * ---------------------------------------------------------------------------------------- */
void or_pack_values(int count, const double *v, int m, float *out)
{
    for (int i = 0; i < count; ++i) {
        uint64_t b;
        memcpy(&b, &v[i], 8);
        int e = (int)((b >> 52) & 0x7ff) - 1023;
        if (e < -126) { out[i] = 0.0f; continue; }
        uint64_t frac = b & ((1ull << 52) - 1), q = frac >> (52 - m);
        uint64_t rem = frac & ((1ull << (52 - m)) - 1), half = 1ull << (51 - m);
        if (rem > half || (rem == half && (q & 1))) ++q;
        if (q >> m) { q = 0; ++e; }
        double mag = ldexp(1.0 + (double)q / (double)(1ull << m), e);
        out[i] = (float)((b >> 63) ? -mag : mag);
    }
}
