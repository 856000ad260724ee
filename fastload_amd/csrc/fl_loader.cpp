// Matrix Market loader: fl_mmio_allinone (replaces mmio_allinone, src/auxiliary.h:65-232, with
// the banner/size parsing of src/mmio.h:142-255).
//
// Semantics kept from the reference:
//   - banner "%%MatrixMarket matrix coordinate <real|complex|pattern|integer>
//     <general|symmetric|hermitian|skew-symmetric>", tokens after the first case-insensitive;
//   - size line = first line not starting with '%';
//   - entries are whitespace-separated numbers (fscanf semantics): pattern -> 1.0,
//     integer -> (double)int, complex -> real part;
//   - symmetric/hermitian files mirror every off-diagonal entry; skew-symmetric is read as
//     general (no mirror), as the reference does;
//   - CSR order: per row, entries in insertion order while scanning the file; the mirror of
//     line (r,c) is inserted into row c when that line is read.
// Added: error codes for truncated files and out-of-range indices (the reference writes out of
// bounds), and the dense "array" layout is refused (the reference cannot read it either).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastload.h"

namespace fl {
void set_error(const char *fmt, ...);
}

namespace {

struct Reader {
    const char *p, *end;
    void skip_ws() {
        while (p < end && isspace((unsigned char)*p)) ++p;
    }
    bool next_long(long &v) {
        skip_ws();
        if (p >= end) return false;
        char *q = nullptr;
        errno = 0;
        v = strtol(p, &q, 10);
        if (q == p || errno) return false;
        p = q;
        return true;
    }
    bool next_double(double &v) {
        skip_ws();
        if (p >= end) return false;
        char *q = nullptr;
        v = strtod(p, &q);
        if (q == p) return false;
        p = q;
        return true;
    }
    // Returns [begin, end) of the next line and advances past its newline.
    bool next_line(const char *&b, const char *&e) {
        if (p >= end) return false;
        b = p;
        while (p < end && *p != '\n') ++p;
        e = p;
        if (p < end) ++p;
        return true;
    }
};

std::string lower(std::string s) {
    for (auto &ch : s) ch = (char)tolower((unsigned char)ch);
    return s;
}

int fail(int code, const char *msg) {
    fl::set_error("mmio_allinone: %s", msg);
    return code;
}

}  // namespace

extern "C" {

void fl_free(void *p) { free(p); }

int fl_mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric, int **csrrowptr, int **csrcolidx, double **csrval,
                     const char *filename) {
    if (!m || !n || !nnz || !isSymmetric || !csrrowptr || !csrcolidx || !csrval || !filename)
        return fail(FL_ERR_INVALID, "NULL argument");
    FILE *f = fopen(filename, "rb");
    if (!f) return fail(FL_ERR_OPEN, "cannot open file");
    std::vector<char> buf;
    {
        char chunk[1 << 16];
        size_t got;
        while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
        fclose(f);
    }
    Reader rd{buf.data(), buf.data() + buf.size()};
    const char *lb, *le;
    if (!rd.next_line(lb, le)) return fail(FL_ERR_BANNER, "empty file");
    char t0[64] = {0}, t1[64] = {0}, t2[64] = {0}, t3[64] = {0}, t4[64] = {0};
    {
        std::string line(lb, le);
        if (sscanf(line.c_str(), "%63s %63s %63s %63s %63s", t0, t1, t2, t3, t4) != 5)
            return fail(FL_ERR_BANNER, "banner needs 5 tokens");
    }
    if (strncmp(t0, "%%MatrixMarket", 14) != 0) return fail(FL_ERR_BANNER, "missing %%MatrixMarket");
    const std::string obj = lower(t1), fmt = lower(t2), field = lower(t3), sym = lower(t4);
    if (obj != "matrix") return fail(FL_ERR_BANNER, "object is not 'matrix'");
    if (fmt == "array") return fail(FL_ERR_UNSUPPORTED, "dense array files are not supported");
    if (fmt != "coordinate") return fail(FL_ERR_BANNER, "unknown storage format");
    int kind;  // 0 real, 1 complex, 2 pattern, 3 integer
    if (field == "real") kind = 0;
    else if (field == "complex") kind = 1;
    else if (field == "pattern") kind = 2;
    else if (field == "integer") kind = 3;
    else return fail(FL_ERR_BANNER, "unknown field type");
    int symm;
    if (sym == "general" || sym == "skew-symmetric") symm = 0;
    else if (sym == "symmetric" || sym == "hermitian") symm = 1;
    else return fail(FL_ERR_BANNER, "unknown symmetry");

    // size line: first line not starting with '%'; if it does not hold three integers, keep
    // reading integers from the following text (mm_read_mtx_crd_size)
    long M = 0, N = 0, NZ = 0;
    for (;;) {
        if (!rd.next_line(lb, le)) return fail(FL_ERR_SIZE, "missing size line");
        if (lb < le && *lb == '%') continue;
        std::string line(lb, le);
        int a, b, c;
        if (sscanf(line.c_str(), "%d %d %d", &a, &b, &c) == 3) {
            M = a; N = b; NZ = c;
        } else {
            if (!rd.next_long(M) || !rd.next_long(N) || !rd.next_long(NZ)) return fail(FL_ERR_SIZE, "bad size line");
        }
        break;
    }
    if (M < 0 || N < 0 || NZ < 0 || M > 0x7fffffffL || N > 0x7fffffffL || NZ > 0x7fffffffL)
        return fail(FL_ERR_SIZE, "invalid sizes");

    std::vector<int> ri(NZ), ci(NZ);
    std::vector<double> vv(NZ);
    std::vector<long long> cnt(M + 1, 0);
    for (long i = 0; i < NZ; ++i) {
        long a, b, iv;
        double fv = 1.0, im;
        bool ok = rd.next_long(a) && rd.next_long(b);
        if (ok) {
            if (kind == 0) ok = rd.next_double(fv);
            else if (kind == 1) ok = rd.next_double(fv) && rd.next_double(im);
            else if (kind == 3) { ok = rd.next_long(iv); fv = (double)iv; }
        }
        if (!ok) return fail(FL_ERR_TRUNC, "truncated entry list");
        if (a < 1 || a > M || b < 1 || b > N) return fail(FL_ERR_RANGE, "entry index out of range");
        ri[i] = (int)(a - 1);
        ci[i] = (int)(b - 1);
        vv[i] = fv;
        cnt[a - 1]++;
    }
    if (symm) {
        for (long i = 0; i < NZ; ++i)
            if (ri[i] != ci[i]) {
                if (ci[i] >= M || ri[i] >= N) return fail(FL_ERR_RANGE, "symmetric mirror outside the matrix");
                cnt[ci[i]]++;
            }
    }
    long long total = 0;
    for (long r = 0; r <= M; ++r) {
        const long long c = cnt[r];
        cnt[r] = total;
        total += c;
    }
    if (total > 0x7fffffffLL) return fail(FL_ERR_SIZE, "more than 2^31-1 entries after mirroring");
    int *rp = (int *)malloc(sizeof(int) * (size_t)(M + 1));
    int *cc = (int *)malloc(sizeof(int) * (size_t)(total ? total : 1));
    double *vo = (double *)malloc(sizeof(double) * (size_t)(total ? total : 1));
    if (!rp || !cc || !vo) {
        free(rp); free(cc); free(vo);
        return fail(FL_ERR_ALLOC, "out of memory");
    }
    std::vector<long long> fill(M + 1, 0);
    for (long r = 0; r <= M; ++r) rp[r] = (int)cnt[r];
    for (long i = 0; i < NZ; ++i) {
        const int r = ri[i], c = ci[i];
        long long off = cnt[r] + fill[r]++;
        cc[off] = c;
        vo[off] = vv[i];
        if (symm && r != c) {
            off = cnt[c] + fill[c]++;
            cc[off] = r;
            vo[off] = vv[i];
        }
    }
    *m = (int)M;
    *n = (int)N;
    *nnz = (int)total;
    *isSymmetric = symm;
    *csrrowptr = rp;
    *csrcolidx = cc;
    *csrval = vo;
    return FL_OK;
}

}  // extern "C"
