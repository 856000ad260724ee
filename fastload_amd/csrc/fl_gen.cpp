// Seeded synthetic CSC generators (include/fastload_gen.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/fastload.h"
#include "../../include/fastload_gen.h"

namespace fl {
void set_error(const char *fmt, ...);
}

namespace {

inline uint64_t splitmix64(uint64_t &z) {
    uint64_t r = (z += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}

// xoshiro256** seeded from (seed, a, b) through splitmix64: an independent stream per column.
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t a, uint64_t b) {
        uint64_t z = seed ^ (a * 0xD1B54A32D192ED03ull) ^ (b * 0x8CB92BA72F3D8DD7ull) ^ 0x5851F42D4C957F2Dull;
        for (auto &w : s) w = splitmix64(z);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint64_t below(uint64_t n) {  // uniform in [0, n), rejection to avoid modulo bias
        if (n <= 1) return 0;
        const uint64_t lim = ~0ull - (~0ull % n);
        uint64_t v;
        do { v = next(); } while (v >= lim);
        return v % n;
    }
    double normal() {  // Box-Muller
        double u1;
        do { u1 = uniform(); } while (u1 <= 0.0);
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    long long poisson(double lam) {
        if (lam < 30.0) {
            const double L = std::exp(-lam);
            long long k = 0;
            double p = 1.0;
            do { ++k; p *= uniform(); } while (p > L);
            return k - 1;
        }
        const double v = std::floor(lam + std::sqrt(lam) * normal() + 0.5);
        return v < 0 ? 0 : (long long)v;
    }
};

// k distinct values of [lo, hi), ascending, appended to out.
void sample_distinct(Rng &r, int lo, int hi, int k, std::vector<int> &out) {
    const long long span = (long long)hi - lo;
    if (k <= 0 || span <= 0) return;
    if (k >= span) {
        for (int v = lo; v < hi; ++v) out.push_back(v);
        return;
    }
    const size_t base = out.size();
    if ((long long)k * 4 > span) {  // selection sampling (Knuth, Algorithm S)
        int need = k;
        for (long long v = 0; v < span && need > 0; ++v)
            if ((double)(span - v) * r.uniform() < need) { out.push_back((int)(lo + v)); --need; }
        return;
    }
    std::vector<int> tmp;
    tmp.reserve(k);
    while ((int)tmp.size() < k) {
        const int want = k - (int)tmp.size();
        for (int i = 0; i < want; ++i) tmp.push_back((int)(lo + (long long)r.below((uint64_t)span)));
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    }
    out.insert(out.begin() + base, tmp.begin(), tmp.end());
}

struct ZipfTable {
    std::vector<double> cdf;
    void build(double s, int kmax) {
        cdf.resize(kmax);
        double acc = 0.0;
        for (int k = 1; k <= kmax; ++k) { acc += std::pow((double)k, -s); cdf[k - 1] = acc; }
        for (auto &c : cdf) c /= acc;
    }
    int draw(Rng &r) const {
        const double u = r.uniform();
        return (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    }
};

int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

template <typename F>
void parallel_for(long long count, F f) {
    const int nt = (int)std::min<long long>(host_threads(), std::max<long long>(count / 4096, 1));
    if (nt <= 1) { f(0LL, count); return; }
    std::vector<std::thread> th;
    const long long chunk = (count + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const long long a = t * chunk, b = std::min(count, a + chunk);
        if (a < b) th.emplace_back([=] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

double param(const double *p, int np, int i, double dflt) { return (p && i < np) ? p[i] : dflt; }

}  // namespace

extern "C" {

int fl_gen_vector(int dist, uint64_t seed, int count, double *out) {
    if (count < 0 || (count && !out)) return FL_ERR_INVALID;
    const long long blocks = (count + 4095) / 4096;
    parallel_for(blocks, [&](long long b0, long long b1) {
        for (long long b = b0; b < b1; ++b) {
            Rng r(seed, (uint64_t)b, 7);
            const long long i1 = std::min<long long>(count, (b + 1) * 4096);
            for (long long i = b * 4096; i < i1; ++i) {
                if (dist == FL_VEC_NORMAL) out[i] = r.normal();
                else if (dist == FL_VEC_UNIFORM) out[i] = 2.0 * r.uniform() - 1.0;
                else out[i] = (double)(i % 10);
            }
        }
    });
    return (dist >= FL_VEC_NORMAL && dist <= FL_VEC_MOD10) ? FL_OK : FL_ERR_INVALID;
}

int fl_gen_matrix(int family, uint64_t seed, int m, int n, const double *params, int np, int **colptr_out,
                  int **rowidx_out, double **val_out, int *nnz_out) {
    if (!colptr_out || !rowidx_out || !val_out || !nnz_out) return FL_ERR_INVALID;
    if (family == FL_GEN_STENCIL27) {
        const long long g = (long long)param(params, np, 0, 128);
        if (g <= 0 || g * g * g > 0x7fffffffLL) return FL_ERR_INVALID;
        m = n = (int)(g * g * g);
    }
    if (m < 0 || n < 0) return FL_ERR_INVALID;
    if (family == FL_GEN_SYMGEOM && m != n) return FL_ERR_INVALID;

    std::vector<long long> cnt((size_t)n + 1, 0);
    ZipfTable zipf;
    const double p0 = param(params, np, 0, 0), p1 = param(params, np, 1, 0), p2 = param(params, np, 2, 0),
                 p3 = param(params, np, 3, 0);

    // --- structurally symmetric twin: built serially (small n)
    std::vector<std::vector<std::pair<int, double>>> sym_cols;
    if (family == FL_GEN_SYMGEOM) {
        const double mean = p0 > 0 ? p0 : 2.6;
        const int cap = p1 > 0 ? (int)p1 : 17;
        const double q = 1.0 / mean;  // geometric on {1,2,...}: mean 1/q
        Rng r(seed, 0x51, 0);
        std::vector<int> stubs;
        for (int i = 0; i < n; ++i) {
            int d;
            do {
                d = 1;
                while (r.uniform() > q) ++d;
            } while (d > cap);
            for (int k = 0; k < d; ++k) stubs.push_back(i);
        }
        for (size_t i = stubs.size(); i > 1; --i) std::swap(stubs[i - 1], stubs[r.below(i)]);
        sym_cols.assign(n, {});
        std::vector<std::vector<int>> adj(n);
        for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
            const int a = stubs[i], b = stubs[i + 1];
            if (a == b) continue;
            if (std::find(adj[a].begin(), adj[a].end(), b) != adj[a].end()) continue;
            adj[a].push_back(b);
            adj[b].push_back(a);
            const double v = 2.0 * r.uniform() - 1.0;
            sym_cols[a].push_back({b, v});
            sym_cols[b].push_back({a, v});
        }
        for (int i = 0; i < n; ++i) {
            sym_cols[i].push_back({i, 2.0 * r.uniform() - 1.0 + 4.0});
            std::sort(sym_cols[i].begin(), sym_cols[i].end(),
                      [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
            cnt[i] = (long long)sym_cols[i].size();
        }
    } else {
        if (family == FL_GEN_ZIPF || family == FL_GEN_ZIPF_DENSE) {
            const int kmax = (int)std::max(1.0, std::min<double>(p1 > 0 ? p1 : 50000, (double)std::max(m, 1)));
            zipf.build(p0 > 0 ? p0 : 1.8423, kmax);
        }
        // per-column counts from stream 0
        const long long G = (family == FL_GEN_STENCIL27) ? (long long)param(params, np, 0, 128) : 0;
        bool bad = false;
        parallel_for(n, [&](long long c0, long long c1) {
            for (long long c = c0; c < c1; ++c) {
                Rng r(seed, (uint64_t)c, 0);
                long long k = 0;
                switch (family) {
                    case FL_GEN_ZIPF: k = m ? zipf.draw(r) : 0; break;
                    case FL_GEN_ZIPF_DENSE:
                        k = (r.uniform() < p2) ? std::min<long long>((long long)p3, m) : (m ? zipf.draw(r) : 0);
                        break;
                    case FL_GEN_STENCIL27: {
                        const long long i = c / (G * G), j = (c / G) % G, l = c % G;
                        auto span = [&](long long v) { return (v > 0) + 1 + (v < G - 1); };
                        k = span(i) * span(j) * span(l);
                        break;
                    }
                    case FL_GEN_BANDED: {
                        const long long h = (long long)(p0 > 0 ? p0 : 5) / 2;
                        const long long lo = std::max(0LL, c - h), hi = std::min<long long>(m, c + h + 1);
                        k = std::max(0LL, hi - lo);
                        break;
                    }
                    case FL_GEN_BLOCKDIAG: {
                        const long long b = (long long)std::max(1.0, p0 > 0 ? p0 : 32);
                        const long long lo = (c / b) * b, hi = std::min<long long>(m, lo + b);
                        const long long blen = std::max(0LL, hi - lo);
                        k = (p1 > 0) ? std::min<long long>((long long)p1, blen) : blen;
                        break;
                    }
                    case FL_GEN_UNIFORM: k = std::min<long long>(r.poisson((double)m * p0), m); break;
                    default: bad = true;
                }
                cnt[c] = k;
            }
        });
        if (bad) return FL_ERR_INVALID;
    }
    long long total = 0;
    for (int c = 0; c < n; ++c) { const long long k = cnt[c]; cnt[c] = total; total += k; }
    cnt[n] = total;
    if (total > 0x7fffffffLL) {
        fl::set_error("fl_gen_matrix: %lld non-zeros exceed int32 indexing", total);
        return FL_ERR_SIZE;
    }
    int *cp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *ri = (int *)malloc(sizeof(int) * (size_t)std::max<long long>(total, 1));
    double *vv = (double *)malloc(sizeof(double) * (size_t)std::max<long long>(total, 1));
    if (!cp || !ri || !vv) {
        free(cp); free(ri); free(vv);
        return FL_ERR_ALLOC;
    }
    for (int c = 0; c <= n; ++c) cp[c] = (int)cnt[c];

    parallel_for(n, [&](long long c0, long long c1) {
        std::vector<int> rows;
        for (long long c = c0; c < c1; ++c) {
            const int base = cp[c], k = cp[c + 1] - cp[c];
            if (family == FL_GEN_SYMGEOM) {
                for (int d = 0; d < k; ++d) { ri[base + d] = sym_cols[c][d].first; vv[base + d] = sym_cols[c][d].second; }
                continue;
            }
            Rng r(seed, (uint64_t)c, 1);
            rows.clear();
            switch (family) {
                case FL_GEN_ZIPF:
                case FL_GEN_ZIPF_DENSE:
                case FL_GEN_UNIFORM: sample_distinct(r, 0, m, k, rows); break;
                case FL_GEN_STENCIL27: {
                    const long long G = (long long)param(params, np, 0, 128);
                    const long long i = c / (G * G), j = (c / G) % G, l = c % G;
                    for (int di = -1; di <= 1; ++di)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int dl = -1; dl <= 1; ++dl) {
                                const long long a = i + di, b = j + dj, e = l + dl;
                                if (a < 0 || b < 0 || e < 0 || a >= G || b >= G || e >= G) continue;
                                rows.push_back((int)((a * G + b) * G + e));
                            }
                    break;
                }
                case FL_GEN_BANDED: {
                    const long long h = (long long)(p0 > 0 ? p0 : 5) / 2;
                    const long long lo = std::max(0LL, c - h);
                    for (int d = 0; d < k; ++d) rows.push_back((int)(lo + d));
                    break;
                }
                case FL_GEN_BLOCKDIAG: {
                    const long long b = (long long)std::max(1.0, p0 > 0 ? p0 : 32);
                    const long long lo = (c / b) * b, hi = std::min<long long>(m, lo + b);
                    sample_distinct(r, (int)lo, (int)hi, k, rows);
                    break;
                }
            }
            for (int d = 0; d < k; ++d) {
                ri[base + d] = rows[d];
                if (family == FL_GEN_STENCIL27) {
                    const double a = 0.5 + r.uniform();
                    vv[base + d] = (r.next() >> 63) ? -a : a;
                } else {
                    vv[base + d] = r.normal();
                }
            }
        }
    });
    *colptr_out = cp;
    *rowidx_out = ri;
    *val_out = vv;
    *nnz_out = (int)total;
    return FL_OK;
}

}  // extern "C"
