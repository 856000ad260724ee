// On-disk cache of the FastLoad preprocessing: a versioned binary file that holds a tiled layout
// (the slide_matrix of src/format.h:6-20 plus the column order) and, for a device-resident matrix,
// every prepared schedule, so later runs skip the Matrix Market parse, the CSR->CSC conversion,
// the column sort, the tiling and the schedule sorts. The reference rebuilds all of this on every
// run (src/main.cu:151-248).
//
// File layout (little-endian, every section 8-byte aligned):
//   FileHeader (fixed 192 bytes): magic "FLMATRIX", version, byte-order mark, sizes, schedule
//     parameters, number of sections
//   SectionEntry[nsections]: tag, element size, element count, byte offset, 64-bit hash
//   section data
// A reader checks the magic, version and byte-order mark, that every section lies inside the
// file and hashes to its recorded value, and that each array has the length the header implies.
// fl_matrix_load then validates every index the kernels will use (row and column ranges,
// pointer arrays, schedule slots and runs) before the matrix is returned, and recomputes what is
// cheap to derive (level table, tile metadata, classification, batch membership, value copies)
// instead of trusting stored copies.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "fl_matrix.h"

namespace fl {

void prepare_values(fl_matrix *A, int kind, hipStream_t s);
int win_install(fl_matrix *A, int bt, bool rows, int nb_file, const uint16_t *slot, const int *win_rows, long long D,
                const int *off, hipStream_t s);
int pb_install(fl_matrix *A, int bt, int R, int nb_file, int C, const int *chunk_batch, const int *chunk_off,
               long long ring, const uint16_t *lslot, const uint16_t *rowlocal, const int *band_start,
               const int *run_off, long long nruns, const int *run_lstart, const int *run_dest, const int *run_len,
               hipStream_t s);
int row_install(fl_matrix *A, const int *row_ptr, const int *row_col, const double *row_val, hipStream_t s);

namespace {

constexpr char kMagic[8] = {'F', 'L', 'M', 'A', 'T', 'R', 'I', 'X'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kByteOrder = 0x01020304u;

enum Tag : uint32_t {
    PERM = 1, SORTNNZ, TILE_PTR, TILE_LEN, TILE_COLIDX, TILE_FORMAT, COUNTROW, SEGOFF, SORTROWIDX, SORTROWINDEX,
    SORTVAL,
    WIN_SLOT = 20, WIN_ROWS, WIN_OFF,
    PB_LSLOT = 30, PB_ROWLOCAL, PB_BAND_START, PB_RUN_OFF, PB_RUN_LSTART, PB_RUN_DEST, PB_RUN_LEN, PB_CHUNK_BATCH,
    PB_CHUNK_OFF,
    ROW_PTR = 40, ROW_COL, ROW_VAL,
};

struct FileHeader {
    char magic[8];
    uint32_t version, byte_order;
    int32_t m, n, nnz, W, levels, tilenum, segsum, max_col_nnz;
    uint32_t value_kinds;  // bit k: the value copy of kind k was prepared (rebuilt on load)
    int32_t device_layout;  // 1: written from a device-resident matrix (fl_matrix_save)
    int32_t bt, ord_kind, nbatches, win_ready, pb_ready, row_ready;
    int32_t pb_R, pb_nbands, pb_chunks, reserved0;
    int64_t pb_ring, pb_nruns, win_D;
    uint32_t nsections, reserved1;
    // HYB schedule: hyb_H1 = heavy columns + 1 when prepared (0: not), with its PB parameters. The
    // schedule is rebuilt on load from the layout (its sorts are redone, about 40 ms at 67 M
    // entries) rather than stored.
    int64_t hyb_H1, hyb_bt, hyb_R;
    int64_t reserved[5];
};
static_assert(sizeof(FileHeader) == 192, "header layout");

struct SectionEntry {
    uint32_t tag, elem_size;
    uint64_t count, offset, hash;
};
static_assert(sizeof(SectionEntry) == 32, "section entry layout");

// 64-bit hash of a byte range, 8 bytes per step (a multiply-rotate mix; detects truncation and
// accidental corruption, not deliberate tampering -- the index checks cover safety).
uint64_t hash_bytes(const void *p, size_t bytes) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    uint64_t h = 0x9e3779b97f4a7c15ull ^ bytes;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, c + i, 8);
        h ^= w * 0xff51afd7ed558ccdull;
        h = (h << 31 | h >> 33) * 0xc4ceb9fe1a85ec53ull;
    }
    uint64_t w = 0;
    if (bytes > i) std::memcpy(&w, c + i, bytes - i);
    h ^= w * 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    return h ^ (h >> 29);
}

struct FileCloser {
    void operator()(FILE *f) const {
        if (f) fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

int io_error(const char *what, const char *path) {
    set_error("%s: %s", what, path);
    return FL_ERR_OPEN;
}

// ------------------------------------------------------------------------------- writer
struct Writer {
    struct Pending {
        uint32_t tag, elem;
        uint64_t count;
        const void *data;
    };
    std::vector<Pending> secs;
    template <typename T>
    void add(uint32_t tag, const T *data, size_t count) {
        secs.push_back({tag, (uint32_t)sizeof(T), (uint64_t)count, data});
    }
    int write(const char *path, FileHeader h) {
        h.nsections = (uint32_t)secs.size();
        std::vector<SectionEntry> table(secs.size());
        uint64_t off = sizeof(FileHeader) + sizeof(SectionEntry) * secs.size();
        for (size_t i = 0; i < secs.size(); ++i) {
            off = (off + 7) & ~7ull;
            const size_t bytes = (size_t)secs[i].count * secs[i].elem;
            table[i] = {secs[i].tag, secs[i].elem, secs[i].count, off, hash_bytes(secs[i].data, bytes)};
            off += bytes;
        }
        File f(fopen(path, "wb"));
        if (!f) return io_error("cannot open for writing", path);
        if (fwrite(&h, sizeof h, 1, f.get()) != 1) return io_error("write failed", path);
        if (!table.empty() && fwrite(table.data(), sizeof(SectionEntry), table.size(), f.get()) != table.size())
            return io_error("write failed", path);
        uint64_t pos = sizeof(FileHeader) + sizeof(SectionEntry) * secs.size();
        static const char zeros[8] = {0};
        for (size_t i = 0; i < secs.size(); ++i) {
            if (table[i].offset > pos && fwrite(zeros, 1, table[i].offset - pos, f.get()) != table[i].offset - pos)
                return io_error("write failed", path);
            const size_t bytes = (size_t)secs[i].count * secs[i].elem;
            if (bytes && fwrite(secs[i].data, 1, bytes, f.get()) != bytes) return io_error("write failed", path);
            pos = table[i].offset + bytes;
        }
        if (fflush(f.get()) != 0) return io_error("write failed", path);
        return FL_OK;
    }
};

FileHeader new_header() {
    FileHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.byte_order = kByteOrder;
    h.ord_kind = -1;
    return h;
}

// ------------------------------------------------------------------------------- reader
struct Reader {
    FileHeader h;
    std::vector<SectionEntry> table;
    std::vector<std::vector<unsigned char>> data;  // per section, read on demand
    File f;
    const char *path = nullptr;
    long long file_size = 0;

    int open(const char *p) {
        path = p;
        f.reset(fopen(p, "rb"));
        if (!f) return io_error("cannot open", p);
        if (fseeko(f.get(), 0, SEEK_END) != 0) return io_error("cannot seek", p);
        file_size = ftello(f.get());
        rewind(f.get());
        if (fread(&h, sizeof h, 1, f.get()) != 1) return bad("file shorter than the header");
        if (std::memcmp(h.magic, kMagic, 8) != 0) return bad("not a FastLoad matrix file (bad magic)");
        if (h.byte_order != kByteOrder) return bad("written with another byte order");
        if (h.version != kVersion) {
            set_error("fl_matrix_load: %s: file version %u, this library reads version %u", p, h.version, kVersion);
            return FL_ERR_INVALID;
        }
        if (h.m < 0 || h.n < 0 || h.nnz < 0 || h.tilenum < 0 || h.segsum < 0 || h.levels < 0 ||
            (h.W != 32 && h.W != 64) || h.nsections > 64)
            return bad("header sizes out of range");
        table.resize(h.nsections);
        if (h.nsections &&
            fread(table.data(), sizeof(SectionEntry), h.nsections, f.get()) != h.nsections)
            return bad("section table truncated");
        for (const SectionEntry &e : table) {
            if (e.elem_size == 0 || e.elem_size > 8 || e.count > (1ull << 40) ||
                e.offset + e.count * e.elem_size > (uint64_t)file_size || e.offset % 8)
                return bad("a section lies outside the file");
        }
        data.resize(h.nsections);
        return FL_OK;
    }
    int bad(const char *what) {
        set_error("fl_matrix_load: %s: %s", path, what);
        return FL_ERR_INVALID;
    }
    const SectionEntry *find(uint32_t tag) const {
        for (const SectionEntry &e : table)
            if (e.tag == tag) return &e;
        return nullptr;
    }
    // Section `tag` as `count` elements of T; nullptr (and an error) when it is missing, has another
    // element size or count, or does not hash to its recorded value.
    template <typename T>
    const T *get(uint32_t tag, uint64_t count, int *rc) {
        const SectionEntry *e = find(tag);
        if (!e || e->elem_size != sizeof(T) || e->count != count) {
            set_error("fl_matrix_load: %s: section %u missing or of the wrong length (%llu, expected %llu)", path, tag,
                      e ? (unsigned long long)e->count : 0ull, (unsigned long long)count);
            *rc = FL_ERR_INVALID;
            return nullptr;
        }
        const size_t i = (size_t)(e - table.data());
        const size_t bytes = (size_t)e->count * e->elem_size;
        std::vector<unsigned char> &buf = data[i];
        buf.resize(std::max<size_t>(bytes, 8));
        if (fseeko(f.get(), (off_t)e->offset, SEEK_SET) != 0 || (bytes && fread(buf.data(), 1, bytes, f.get()) != bytes)) {
            *rc = io_error("read failed", path);
            return nullptr;
        }
        if (hash_bytes(buf.data(), bytes) != e->hash) {
            set_error("fl_matrix_load: %s: section %u does not match its hash (damaged file)", path, tag);
            *rc = FL_ERR_INVALID;
            return nullptr;
        }
        return reinterpret_cast<const T *>(buf.data());
    }
    void drop(uint32_t tag) {
        if (const SectionEntry *e = find(tag)) std::vector<unsigned char>().swap(data[(size_t)(e - table.data())]);
    }
};

template <typename T>
std::vector<T> to_host(const DevBuf<T> &d, size_t count, hipStream_t s) {
    std::vector<T> h(std::max<size_t>(count, 1));
    if (count) FL_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    return h;
}

template <typename T>
T *malloc_copy(const T *src, size_t count) {
    T *p = (T *)malloc(sizeof(T) * std::max<size_t>(count, 1));
    if (!p) throw std::bad_alloc();
    if (count) std::memcpy(p, src, sizeof(T) * count);
    return p;
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

int fl_slide_save(const char *path, const slide_matrix *M, int m, int n, int nnz, int h_count, const int *perm) {
    FL_API_BEGIN
    if (!path || !M || m < 0 || n < 0 || nnz < 0 || M->tilenum < 0 || M->segsum < 0) {
        set_error("fl_slide_save: NULL or negative argument");
        return FL_ERR_INVALID;
    }
    const int T = M->tilenum;
    if ((T > 0 && (!M->tile_ptr || !M->tile_len || !M->tile_colidx || !M->tile_format || !M->countrow)) ||
        (nnz > 0 && (!M->sortrowidx || !M->sortval || !M->sortrowindex)) || (M->segsum > 0 && !M->segmentoffset)) {
        set_error("fl_slide_save: slide_matrix arrays missing");
        return FL_ERR_INVALID;
    }
    int maxlen = 0;
    for (int t = 0; t < T; ++t) maxlen = std::max(maxlen, M->tile_len[t]);
    FileHeader h = new_header();
    h.m = m;
    h.n = n;
    h.nnz = nnz;
    h.W = maxlen > 32 ? 64 : 32;
    h.levels = h_count;
    h.tilenum = T;
    h.segsum = M->segsum;
    Writer w;
    if (perm) w.add(PERM, perm, n);
    const int zero = 0;
    w.add(TILE_PTR, T > 0 ? M->tile_ptr : &zero, (size_t)T + 1);
    w.add(TILE_LEN, M->tile_len, T);
    w.add(TILE_COLIDX, M->tile_colidx, T);
    w.add(TILE_FORMAT, M->tile_format, T);
    w.add(COUNTROW, T > 0 ? M->countrow : &zero, (size_t)T + 1);
    w.add(SEGOFF, M->segsum > 0 ? M->segmentoffset : &zero, (size_t)M->segsum + 1);
    w.add(SORTROWIDX, M->sortrowidx, nnz);
    w.add(SORTROWINDEX, M->sortrowindex, nnz);
    w.add(SORTVAL, M->sortval, nnz);
    return w.write(path, h);
    FL_API_END
}

int fl_slide_load(const char *path, slide_matrix *M, int *m, int *n, int *nnz, int *h_count, int **perm) {
    FL_API_BEGIN
    if (!path || !M || !m || !n || !nnz) {
        set_error("fl_slide_load: NULL argument");
        return FL_ERR_INVALID;
    }
    std::memset(M, 0, sizeof *M);
    if (perm) *perm = nullptr;
    Reader r;
    int rc = r.open(path);
    if (rc) return rc;
    const FileHeader &h = r.h;
    const size_t T = h.tilenum, nz = h.nnz, S = h.segsum;
    const int *tp = r.get<int>(TILE_PTR, T + 1, &rc);
    const int *tl = tp ? r.get<int>(TILE_LEN, T, &rc) : nullptr;
    const int *tc = tl ? r.get<int>(TILE_COLIDX, T, &rc) : nullptr;
    const int *tf = tc ? r.get<int>(TILE_FORMAT, T, &rc) : nullptr;
    const int *cr = tf ? r.get<int>(COUNTROW, T + 1, &rc) : nullptr;
    const int *so = cr ? r.get<int>(SEGOFF, S + 1, &rc) : nullptr;
    const int *ri = so ? r.get<int>(SORTROWIDX, nz, &rc) : nullptr;
    const int *rx = ri ? r.get<int>(SORTROWINDEX, nz, &rc) : nullptr;
    const double *sv = rx ? r.get<double>(SORTVAL, nz, &rc) : nullptr;
    if (!sv) return rc;
    const int *pm = nullptr;
    if (perm && r.find(PERM)) {
        pm = r.get<int>(PERM, (uint64_t)h.n, &rc);
        if (!pm) return rc;
    }
    slide_matrix out;
    std::memset(&out, 0, sizeof out);
    try {
        out.tile_ptr = malloc_copy(tp, T + 1);
        out.tile_len = malloc_copy(tl, T);
        out.tile_colidx = malloc_copy(tc, T);
        out.tile_format = malloc_copy(tf, T);
        out.countrow = malloc_copy(cr, T + 1);
        out.segmentoffset = malloc_copy(so, S + 1);
        out.sortrowidx = malloc_copy(ri, nz);
        out.sortrowindex = malloc_copy(rx, nz);
        out.sortval = malloc_copy(sv, nz);
        if (pm) *perm = malloc_copy(pm, (size_t)h.n);
    } catch (...) {
        fl_Tile_destroy(&out);
        throw;
    }
    out.tilenum = (int)T;
    out.segsum = (int)S;
    *M = out;
    *m = h.m;
    *n = h.n;
    *nnz = h.nnz;
    if (h_count) *h_count = h.levels;
    return FL_OK;
    FL_API_END
}

int fl_matrix_save(const fl_matrix *A, const char *path, void *stream) {
    FL_API_BEGIN
    if (!A || !path) {
        set_error("fl_matrix_save: NULL argument");
        return FL_ERR_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t T = A->tilenum, nz = A->nnz, n = A->n, m = A->m, S = A->segsum;
    FileHeader h = new_header();
    h.m = A->m;
    h.n = A->n;
    h.nnz = A->nnz;
    h.W = A->W;
    h.levels = A->L;
    h.tilenum = A->tilenum;
    h.segsum = A->segsum;
    h.max_col_nnz = A->max_col_nnz;
    h.device_layout = 1;
    h.value_kinds = (A->val32.p ? 1u << FL_VAL_F32 : 0u) | (A->bf16_hi.p ? 1u << FL_VAL_BF16 : 0u) |
                    (A->bf16_pair.p ? 1u << FL_VAL_BF16X2 : 0u) | (A->pack_wanted ? 1u << FL_VAL_PACKED : 0u);
    auto perm = to_host(A->perm, n, s), sortnnz = to_host(A->sortnnz, n, s);
    auto tp = to_host(A->tile_ptr, T + 1, s), tl = to_host(A->tile_len, T, s), tc = to_host(A->tile_colidx, T, s),
         tf = to_host(A->tile_format, T, s), cr = to_host(A->countrow, T + 1, s),
         so = to_host(A->segmentoffset, S + 1, s);
    auto ri = to_host(A->sortrowidx, nz, s), rx = to_host(A->sortrowindex, nz, s);
    auto sv = to_host(A->val64, nz, s);
    Writer w;
    w.add(PERM, perm.data(), n);
    w.add(SORTNNZ, sortnnz.data(), n);
    w.add(TILE_PTR, tp.data(), T + 1);
    w.add(TILE_LEN, tl.data(), T);
    w.add(TILE_COLIDX, tc.data(), T);
    w.add(TILE_FORMAT, tf.data(), T);
    w.add(COUNTROW, cr.data(), T + 1);
    w.add(SEGOFF, so.data(), S + 1);
    w.add(SORTROWIDX, ri.data(), nz);
    w.add(SORTROWINDEX, rx.data(), nz);
    w.add(SORTVAL, sv.data(), nz);
    std::vector<uint16_t> wslot, lslot, rowlocal;
    std::vector<int> wrows, woff, band_start, run_off, lstart, dest, len, rptr, rcol;
    std::vector<double> rval;
    if (A->win_ready || A->pb_ready) {
        h.bt = A->bt;
        h.ord_kind = A->ord_kind;
        h.nbatches = A->nbatches;
    }
    if (A->win_ready) {
        const size_t nb = A->nbatches;
        woff = to_host(A->win_off, nb + 1, s);
        FL_HIP(hipStreamSynchronize(s));
        const size_t D = (size_t)woff[nb];
        h.win_ready = 1;
        h.win_D = (int64_t)D;
        wslot = to_host(A->win_slot, nz, s);
        wrows = to_host(A->win_rows, D, s);
        w.add(WIN_SLOT, wslot.data(), nz);
        w.add(WIN_ROWS, wrows.data(), D);
        w.add(WIN_OFF, woff.data(), nb + 1);
    }
    if (A->pb_ready) {
        const size_t nb = A->nbatches, C = A->pb_chunks, nr = A->nruns;
        h.pb_ready = 1;
        h.pb_R = A->pb_R;
        h.pb_nbands = A->nbands;
        h.pb_chunks = A->pb_chunks;
        h.pb_ring = A->pb_ring;
        h.pb_nruns = A->nruns;
        lslot = to_host(A->pb_lslot, nz, s);
        rowlocal = to_host(A->pb_rowlocal, nz, s);
        band_start = to_host(A->pb_band_start, C * A->nbands + 1, s);
        run_off = to_host(A->pb_run_off, nb + 1, s);
        lstart = to_host(A->pb_run_lstart, nr, s);
        dest = to_host(A->pb_run_dest, nr, s);
        len = to_host(A->pb_run_len, nr, s);
        w.add(PB_LSLOT, lslot.data(), nz);
        w.add(PB_ROWLOCAL, rowlocal.data(), nz);
        w.add(PB_BAND_START, band_start.data(), C * A->nbands + 1);
        w.add(PB_RUN_OFF, run_off.data(), nb + 1);
        w.add(PB_RUN_LSTART, lstart.data(), nr);
        w.add(PB_RUN_DEST, dest.data(), nr);
        w.add(PB_RUN_LEN, len.data(), nr);
        w.add(PB_CHUNK_BATCH, A->pb_chunk_batch.data(), C + 1);
        w.add(PB_CHUNK_OFF, A->pb_chunk_off.data(), C);
    }
    if (A->hyb_ready) {
        h.hyb_H1 = (int64_t)A->hyb_H + 1;
        h.hyb_bt = A->hyb_bt;
        h.hyb_R = A->hyb_R;
    }
    if (A->row_ready) {
        h.row_ready = 1;
        rptr = to_host(A->row_ptr, m + 1, s);
        rcol = to_host(A->row_col, nz, s);
        rval = to_host(A->row_val64, nz, s);
        w.add(ROW_PTR, rptr.data(), m + 1);
        w.add(ROW_COL, rcol.data(), nz);
        w.add(ROW_VAL, rval.data(), nz);
    }
    FL_HIP(hipStreamSynchronize(s));
    return w.write(path, h);
    FL_API_END
}

int fl_matrix_load(fl_matrix **out, const char *path, void *stream) {
    FL_API_BEGIN
    if (!out || !path) {
        set_error("fl_matrix_load: NULL argument");
        return FL_ERR_INVALID;
    }
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    Reader r;
    int rc = r.open(path);
    if (rc) return rc;
    const FileHeader &h = r.h;
    const int m = h.m, n = h.n, nnz = h.nnz;
    // the column order and the elements in tile order; everything else is recomputed or checked
    const int *perm = r.get<int>(PERM, (uint64_t)n, &rc);
    const int *sortnnz = perm ? r.get<int>(SORTNNZ, (uint64_t)n, &rc) : nullptr;
    const int *rows = sortnnz ? r.get<int>(SORTROWIDX, (uint64_t)nnz, &rc) : nullptr;
    const double *vals = rows ? r.get<double>(SORTVAL, (uint64_t)nnz, &rc) : nullptr;
    if (!vals) return rc;
    {
        std::vector<char> seen((size_t)std::max(n, 1), 0);
        for (int i = 0; i < n; ++i) {
            if (perm[i] < 0 || perm[i] >= n || seen[perm[i]]) return r.bad("column order is not a permutation");
            seen[perm[i]] = 1;
        }
        long long sum = 0;
        for (int i = 0; i < n; ++i) {
            if (sortnnz[i] < 0 || sortnnz[i] > m || (i && sortnnz[i] < sortnnz[i - 1]))
                return r.bad("sorted column lengths not ascending within [0, m]");
            sum += sortnnz[i];
        }
        if (sum != nnz) return r.bad("sorted column lengths do not add up to nnz");
        for (int i = 0; i < nnz; ++i)
            if (rows[i] < 0 || rows[i] >= m) return r.bad("row index out of range");
    }
    std::unique_ptr<fl_matrix> A(new fl_matrix());
    A->m = m;
    A->n = n;
    A->nnz = nnz;
    A->W = h.W;
    A->perm.alloc(std::max(n, 1));
    A->sortnnz.alloc(std::max(n, 1));
    A->sortrowidx.alloc(std::max(nnz, 1));
    A->val64.alloc(std::max(nnz, 1));
    if (n) {
        FL_HIP(hipMemcpyAsync(A->perm.p, perm, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(A->sortnnz.p, sortnnz, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    }
    if (nnz) {
        FL_HIP(hipMemcpyAsync(A->sortrowidx.p, rows, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(A->val64.p, vals, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, s));
    }
    A->max_col_nnz = n ? sortnnz[n - 1] : 0;
    FL_HIP(hipStreamSynchronize(s));
    r.drop(SORTROWIDX);
    r.drop(SORTVAL);
    ElemSource src;
    src.tiled = true;
    transform_device(A.get(), A->sortnnz.p, src, s, nullptr, nullptr);
    // the recomputed tile layout must be the stored one (a file of another layout is refused)
    if (A->tilenum != h.tilenum || A->L != h.levels || A->segsum != h.segsum) return r.bad("tile layout differs");
    {
        const size_t T = A->tilenum;
        const int *tp = r.get<int>(TILE_PTR, T + 1, &rc);
        const int *tf = tp ? r.get<int>(TILE_FORMAT, T, &rc) : nullptr;
        const int *rx = tf ? r.get<int>(SORTROWINDEX, (uint64_t)nnz, &rc) : nullptr;
        if (!rx) return rc;
        auto dtp = to_host(A->tile_ptr, T + 1, s);
        auto dtf = to_host(A->tile_format, T, s);
        auto drx = to_host(A->sortrowindex, (size_t)nnz, s);
        FL_HIP(hipStreamSynchronize(s));
        if (std::memcmp(dtp.data(), tp, sizeof(int) * (T + 1)) || (T && std::memcmp(dtf.data(), tf, sizeof(int) * T)) ||
            (nnz && std::memcmp(drx.data(), rx, sizeof(int) * (size_t)nnz)))
            return r.bad("stored tile metadata differs from the layout of the stored elements");
        r.drop(SORTROWINDEX);
    }
    for (int k = FL_VAL_F32; k <= FL_VAL_BF16X2; ++k)
        if (h.value_kinds & (1u << k)) prepare_values(A.get(), k, s);
    if (h.win_ready || h.pb_ready) {
        if (A->W != 64 || (h.bt != 64 && h.bt != 128 && h.bt != 256) || (h.ord_kind != 0 && h.ord_kind != 1) ||
            (h.pb_ready && h.ord_kind != 0) || h.nbatches < 0)
            return r.bad("schedule parameters out of range");
    }
    if (h.win_ready) {
        const uint64_t nb = (uint64_t)h.nbatches;
        if (h.win_D < 0 || h.win_D > nnz) return r.bad("WIN slot count out of range");
        const uint16_t *slot = r.get<uint16_t>(WIN_SLOT, (uint64_t)nnz, &rc);
        const int *wrows = slot ? r.get<int>(WIN_ROWS, (uint64_t)h.win_D, &rc) : nullptr;
        const int *woff = wrows ? r.get<int>(WIN_OFF, nb + 1, &rc) : nullptr;
        if (!woff) return rc;
        rc = win_install(A.get(), h.bt, h.ord_kind == 1, h.nbatches, slot, wrows, h.win_D, woff, s);
        if (rc) return rc;
    }
    if (h.pb_ready) {
        const uint64_t nb = (uint64_t)h.nbatches, C = (uint64_t)std::max(h.pb_chunks, 0);
        if (h.pb_chunks < 1 || h.pb_chunks > 64 || h.pb_nbands < 1 || h.pb_nruns < 0 || h.pb_nruns > nnz)
            return r.bad("PB parameters out of range");
        const uint16_t *lslot = r.get<uint16_t>(PB_LSLOT, (uint64_t)nnz, &rc);
        const uint16_t *rowlocal = lslot ? r.get<uint16_t>(PB_ROWLOCAL, (uint64_t)nnz, &rc) : nullptr;
        const int *bstart = rowlocal ? r.get<int>(PB_BAND_START, C * (uint64_t)h.pb_nbands + 1, &rc) : nullptr;
        const int *roff = bstart ? r.get<int>(PB_RUN_OFF, nb + 1, &rc) : nullptr;
        const int *rls = roff ? r.get<int>(PB_RUN_LSTART, (uint64_t)h.pb_nruns, &rc) : nullptr;
        const int *rde = rls ? r.get<int>(PB_RUN_DEST, (uint64_t)h.pb_nruns, &rc) : nullptr;
        const int *rle = rde ? r.get<int>(PB_RUN_LEN, (uint64_t)h.pb_nruns, &rc) : nullptr;
        const int *cb = rle ? r.get<int>(PB_CHUNK_BATCH, C + 1, &rc) : nullptr;
        const int *co = cb ? r.get<int>(PB_CHUNK_OFF, C, &rc) : nullptr;
        if (!co) return rc;
        if (h.pb_nbands != std::max((m + h.pb_R - 1) / std::max(h.pb_R, 1), 1)) return r.bad("PB band count");
        rc = pb_install(A.get(), h.bt, h.pb_R, h.nbatches, h.pb_chunks, cb, co, h.pb_ring, lslot, rowlocal, bstart,
                        roff, h.pb_nruns, rls, rde, rle, s);
        if (rc) return rc;
    }
    if (h.row_ready) {
        const int *rp = r.get<int>(ROW_PTR, (uint64_t)m + 1, &rc);
        const int *rcol = rp ? r.get<int>(ROW_COL, (uint64_t)nnz, &rc) : nullptr;
        const double *rv = rcol ? r.get<double>(ROW_VAL, (uint64_t)nnz, &rc) : nullptr;
        if (!rv) return rc;
        rc = row_install(A.get(), rp, rcol, rv, s);
        if (rc) return rc;
    }
    if (h.value_kinds & (1u << FL_VAL_PACKED)) prepare_values(A.get(), FL_VAL_PACKED, s);  // after the schedules
    if (h.hyb_H1 != 0) {
        const int64_t H = h.hyb_H1 - 1;
        if (H < 0 || H > std::min<int64_t>(n, kHybMaxH32) || (h.hyb_bt != 64 && h.hyb_bt != 128 && h.hyb_bt != 256) ||
            h.hyb_R < 1024 || h.hyb_R > 16384)
            return r.bad("HYB schedule parameters out of range");
        hyb_prepare(A.get(), (int)H, (int)h.hyb_bt, (int)h.hyb_R, s);
    }
    FL_HIP(hipStreamSynchronize(s));
    *out = A.release();
    return FL_OK;
    FL_API_END
}

}  // extern "C"
