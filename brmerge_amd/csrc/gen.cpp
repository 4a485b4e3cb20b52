// Deterministic synthetic CSR generators, MatrixMarket I/O and a host CSR
// transpose (include/brmerge_gen.h).  Host C++ only; parallelised with
// std::thread over independent, counter-seeded random streams.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/brmerge.h"
#include "../../include/brmerge_gen.h"

namespace {

// ---------------------------------------------------------------- RNG
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

enum Domain : uint64_t { DOM_ER = 1, DOM_RMAT = 2, DOM_VAL = 3, DOM_BAND = 4, DOM_ERV = 5 };

// xoshiro256** seeded by SplitMix64 from (seed, domain, stream id).
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t domain, uint64_t id) {
        uint64_t x = seed ^ (domain << 56) ^ (id * 0xD1B54A32D192ED03ull);
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
    double sym() { return 2.0 * u01() - 1.0; }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

// counter-based value for global entry k
inline double value_at(uint64_t seed, uint64_t k) {
    uint64_t x = seed ^ (DOM_VAL << 56) ^ (k * 0xD1B54A32D192ED03ull);
    const uint64_t z = splitmix64(x);
    return 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
}

int nthreads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 4;
    return (int)std::min(h, 32u);
}

template <typename F>
void parallel_for(int64_t n, F f) {
    const int T = nthreads();
    if (n < 4096 || T <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=]() { f(b, e); });
    }
    for (auto& x : th) x.join();
}

template <typename T>
T* new_arr(int64_t n) {
    return new (std::nothrow) T[(size_t)(n > 0 ? n : 1)];
}

// Builds CSR from per-row counts and a row filler: fill(i, col_out, val_out).
template <typename Fill>
int build_csr(int M, const std::vector<int>& cnt, Fill fill, int** rpt_, int** col_, double** val_) {
    int64_t nnz = 0;
    for (int i = 0; i < M; ++i) nnz += cnt[i];
    if (nnz > INT32_MAX) return BRM_E_OVERFLOW;
    int* rpt = new_arr<int>((int64_t)M + 1);
    int* col = new_arr<int>(nnz);
    double* val = new_arr<double>(nnz);
    if (!rpt || !col || !val) {
        delete[] rpt;
        delete[] col;
        delete[] val;
        return BRM_E_NOMEM;
    }
    rpt[0] = 0;
    for (int i = 0; i < M; ++i) rpt[i + 1] = rpt[i] + cnt[i];
    parallel_for(M, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) fill((int)i, col + rpt[i], val + rpt[i]);
    });
    *rpt_ = rpt;
    *col_ = col;
    *val_ = val;
    return BRM_OK;
}

// Floyd's algorithm: d distinct values of [0, n), sorted.
void floyd(Rng& r, int64_t n, int d, std::vector<int64_t>& out) {
    out.clear();
    for (int64_t j = n - d; j < n; ++j) {
        const int64_t t = (int64_t)r.below((uint64_t)j + 1);
        if (std::find(out.begin(), out.end(), t) == out.end()) out.push_back(t);
        else out.push_back(j);
    }
    std::sort(out.begin(), out.end());
}

}  // namespace

extern "C" {

int brm_gen_er(int M, int N, int d, uint64_t seed, int** rpt, int** col, double** val) {
    return brm_gen_er_var(M, N, d, d, seed, rpt, col, val);
}

int brm_gen_er_var(int M, int N, int dmin, int dmax, uint64_t seed, int** rpt, int** col, double** val) {
    if (M < 0 || N < 0 || dmin < 0 || dmax < dmin || !rpt || !col || !val) return BRM_E_INVALID;
    const bool fixed = dmin == dmax;
    std::vector<int> cnt(M);
    for (int i = 0; i < M; ++i) {
        int d = dmin;
        if (!fixed) {
            Rng r(seed, DOM_ERV, (uint64_t)i);
            d = dmin + (int)r.below((uint64_t)(dmax - dmin + 1));
        }
        cnt[i] = std::min(d, N);
    }
    return build_csr(M, cnt, [&](int i, int* c, double* v) {
        Rng r(seed, DOM_ER, (uint64_t)i);
        std::vector<int64_t> s;
        floyd(r, N, cnt[i], s);
        for (int k = 0; k < cnt[i]; ++k) c[k] = (int)s[k];
        for (int k = 0; k < cnt[i]; ++k) v[k] = r.sym();
    }, rpt, col, val);
}

int brm_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int** rpt_, int** col_,
                 double** val_) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || !rpt_ || !col_ || !val_) return BRM_E_INVALID;
    const int64_t n = (int64_t)1 << scale;
    const int64_t E = (int64_t)edge_factor * n;
    std::vector<uint32_t> er(E), ec(E);
    const double ab = a + b, abc = a + b + c;
    parallel_for(E, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            Rng r(seed, DOM_RMAT, (uint64_t)e);
            uint32_t row = 0, cc = 0;
            for (int l = 0; l < scale; ++l) {
                const double u = r.u01();
                const uint32_t rb = u >= ab;             // quadrants c, d
                const uint32_t cb = (u >= a && u < ab) || u >= abc;  // quadrants b, d
                row = (row << 1) | rb;
                cc = (cc << 1) | cb;
            }
            er[e] = row;
            ec[e] = cc;
        }
    });
    // bucket by row, then sort + unique each row
    std::vector<std::atomic<int>> hist((size_t)n);
    for (auto& h : hist) h.store(0, std::memory_order_relaxed);
    parallel_for(E, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) hist[er[e]].fetch_add(1, std::memory_order_relaxed);
    });
    std::vector<int64_t> start((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; ++i) start[i + 1] = start[i] + hist[i].load(std::memory_order_relaxed);
    std::vector<std::atomic<int64_t>> pos((size_t)n);
    for (int64_t i = 0; i < n; ++i) pos[i].store(start[i], std::memory_order_relaxed);
    std::vector<uint32_t> bucket(E);
    parallel_for(E, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) bucket[pos[er[e]].fetch_add(1, std::memory_order_relaxed)] = ec[e];
    });
    std::vector<int> cnt((size_t)n);
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            uint32_t* b0 = bucket.data() + start[i];
            uint32_t* b1 = bucket.data() + start[i + 1];
            std::sort(b0, b1);
            cnt[i] = (int)(std::unique(b0, b1) - b0);
        }
    });
    int st = build_csr((int)n, cnt, [&](int i, int* c, double*) {
        const uint32_t* b0 = bucket.data() + start[i];
        for (int k = 0; k < cnt[i]; ++k) c[k] = (int)b0[k];
    }, rpt_, col_, val_);
    if (st != BRM_OK) return st;
    int* rpt = *rpt_;
    double* val = *val_;
    parallel_for((int64_t)rpt[n], [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) val[k] = value_at(seed, (uint64_t)k);
    });
    return BRM_OK;
}

int brm_gen_stencil27(int n, int** rpt, int** col, double** val) {
    if (n < 1 || (int64_t)n * n * n > INT32_MAX || !rpt || !col || !val) return BRM_E_INVALID;
    const int64_t M = (int64_t)n * n * n;
    auto c1 = [n](int x) { return 1 + (x > 0) + (x < n - 1); };
    std::vector<int> cnt((size_t)M);
    for (int64_t r = 0; r < M; ++r) {
        const int x = (int)(r % n), y = (int)((r / n) % n), z = (int)(r / ((int64_t)n * n));
        cnt[r] = c1(x) * c1(y) * c1(z);
    }
    return build_csr((int)M, cnt, [&](int r, int* c, double* v) {
        const int x = r % n, y = (r / n) % n, z = r / (n * n);
        int k = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx, yy = y + dy, zz = z + dz;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
                    c[k] = (zz * n + yy) * n + xx;
                    v[k] = (dx == 0 && dy == 0 && dz == 0) ? 26.0 : -1.0;
                    ++k;
                }
    }, rpt, col, val);
}

int brm_gen_aggregation(int n, int g, int** rpt, int** col, double** val) {
    if (n < 1 || g < 1 || (int64_t)n * n * n > INT32_MAX || !rpt || !col || !val) return BRM_E_INVALID;
    const int64_t M = (int64_t)n * n * n;
    const int m = (n + g - 1) / g;
    std::vector<int> cnt((size_t)M, 1);
    return build_csr((int)M, cnt, [&](int r, int* c, double* v) {
        const int x = r % n, y = (r / n) % n, z = r / (n * n);
        c[0] = ((z / g) * m + y / g) * m + x / g;
        v[0] = 1.0;
    }, rpt, col, val);
}

int brm_gen_banded(int M, int bw, int dmin, int dmax, uint64_t seed, int** rpt, int** col, double** val) {
    if (M < 0 || bw < 0 || dmin < 0 || dmax < dmin || !rpt || !col || !val) return BRM_E_INVALID;
    std::vector<int> cnt(M);
    for (int i = 0; i < M; ++i) {
        Rng r(seed, DOM_BAND, (uint64_t)i);
        const int lo = std::max(0, i - bw), hi = std::min(M - 1, i + bw);
        const int d = dmin + (int)r.below((uint64_t)(dmax - dmin + 1));
        cnt[i] = std::min(d, hi - lo + 1);
    }
    return build_csr(M, cnt, [&](int i, int* c, double* v) {
        Rng r(seed, DOM_BAND, (uint64_t)i);
        (void)r.below((uint64_t)(dmax - dmin + 1));  // the count draw
        const int lo = std::max(0, i - bw), hi = std::min(M - 1, i + bw);
        std::vector<int64_t> s;
        floyd(r, hi - lo + 1, cnt[i], s);
        for (int k = 0; k < cnt[i]; ++k) c[k] = lo + (int)s[k];
        for (int k = 0; k < cnt[i]; ++k) v[k] = r.sym();
    }, rpt, col, val);
}

int brm_csr_transpose(int M, int N, const int* rpt, const int* col, const double* val, int** trpt_, int** tcol_,
                      double** tval_) {
    if (M < 0 || N < 0 || !rpt || !trpt_ || !tcol_ || !tval_) return BRM_E_INVALID;
    const int64_t nnz = rpt[M];
    int* trpt = new_arr<int>((int64_t)N + 1);
    int* tcol = new_arr<int>(nnz);
    double* tval = new_arr<double>(nnz);
    if (!trpt || !tcol || !tval) {
        delete[] trpt;
        delete[] tcol;
        delete[] tval;
        return BRM_E_NOMEM;
    }
    std::fill(trpt, trpt + N + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) {
        if (col[k] < 0 || col[k] >= N) {
            delete[] trpt;
            delete[] tcol;
            delete[] tval;
            return BRM_E_INVALID;
        }
        trpt[col[k] + 1]++;
    }
    for (int j = 0; j < N; ++j) trpt[j + 1] += trpt[j];
    std::vector<int> next(trpt, trpt + N);
    for (int i = 0; i < M; ++i)
        for (int k = rpt[i]; k < rpt[i + 1]; ++k) {
            const int p = next[col[k]]++;
            tcol[p] = i;
            tval[p] = val ? val[k] : 1.0;
        }
    *trpt_ = trpt;
    *tcol_ = tcol;
    *tval_ = tval;
    return BRM_OK;
}

int brm_read_mtx(const char* path, int* M_, int* N_, int64_t* nnz_, int** rpt_, int** col_, double** val_) {
    if (!path || !M_ || !N_ || !nnz_ || !rpt_ || !col_ || !val_) return BRM_E_INVALID;
    std::ifstream in(path);
    if (!in) return BRM_E_INVALID;
    std::string line;
    if (!std::getline(in, line)) return BRM_E_INVALID;
    std::istringstream hs(line);
    std::string tag, obj, fmt, type, sym;
    hs >> tag >> obj >> fmt >> type >> sym;
    auto lower = [](std::string s) {
        for (auto& ch : s) ch = (char)tolower(ch);
        return s;
    };
    obj = lower(obj);
    fmt = lower(fmt);
    type = lower(type);
    sym = lower(sym);
    if (tag != "%%MatrixMarket" || obj != "matrix" || fmt != "coordinate") return BRM_E_INVALID;
    if (type != "real" && type != "integer" && type != "pattern" && type != "complex") return BRM_E_INVALID;
    if (sym != "general" && sym != "symmetric" && sym != "skew-symmetric") return BRM_E_INVALID;
    do {
        if (!std::getline(in, line)) return BRM_E_INVALID;
    } while (line.empty() || line[0] == '%');
    long long M = 0, N = 0, nz = 0;
    {
        std::istringstream ds(line);
        if (!(ds >> M >> N >> nz) || M < 0 || N < 0 || nz < 0 || M > INT32_MAX || N > INT32_MAX) return BRM_E_INVALID;
    }
    std::vector<long long> I, J;
    std::vector<double> V;
    I.reserve((size_t)nz);
    J.reserve((size_t)nz);
    V.reserve((size_t)nz);
    for (long long k = 0; k < nz; ++k) {
        long long i, j;
        double v = 1.0, im;
        if (!(in >> i >> j)) return BRM_E_INVALID;
        if (type == "complex") {
            if (!(in >> v >> im)) return BRM_E_INVALID;
        } else if (type != "pattern") {
            if (!(in >> v)) return BRM_E_INVALID;
        }
        --i;
        --j;
        if (i < 0 || j < 0 || i >= M || j >= N) return BRM_E_INVALID;
        I.push_back(i);
        J.push_back(j);
        V.push_back(v);
        if (sym != "general" && i != j) {
            I.push_back(j);
            J.push_back(i);
            V.push_back(sym == "skew-symmetric" ? -v : v);
        }
    }
    const size_t n = I.size();
    if ((long long)n > INT32_MAX) return BRM_E_OVERFLOW;
    std::vector<size_t> ord(n);
    for (size_t k = 0; k < n; ++k) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
        return I[x] != I[y] ? I[x] < I[y] : J[x] < J[y];
    });
    int* rpt = new_arr<int>(M + 1);
    int* col = new_arr<int>((int64_t)n);
    double* val = new_arr<double>((int64_t)n);
    if (!rpt || !col || !val) {
        delete[] rpt;
        delete[] col;
        delete[] val;
        return BRM_E_NOMEM;
    }
    std::fill(rpt, rpt + M + 1, 0);
    for (size_t k = 0; k < n; ++k) {
        const size_t s = ord[k];
        if (k > 0 && I[ord[k - 1]] == I[s] && J[ord[k - 1]] == J[s]) {  // duplicate entry
            delete[] rpt;
            delete[] col;
            delete[] val;
            return BRM_E_INVALID;
        }
        rpt[I[s] + 1]++;
        col[k] = (int)J[s];
        val[k] = V[s];
    }
    for (long long i = 0; i < M; ++i) rpt[i + 1] += rpt[i];
    *M_ = (int)M;
    *N_ = (int)N;
    *nnz_ = (int64_t)n;
    *rpt_ = rpt;
    *col_ = col;
    *val_ = val;
    return BRM_OK;
}

int brm_write_mtx(const char* path, int M, int N, const int* rpt, const int* col, const double* val) {
    if (!path || M < 0 || N < 0 || !rpt) return BRM_E_INVALID;
    FILE* f = fopen(path, "w");
    if (!f) return BRM_E_INVALID;
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", M, N, rpt[M]);
    for (int i = 0; i < M; ++i)
        for (int k = rpt[i]; k < rpt[i + 1]; ++k) fprintf(f, "%d %d %.17g\n", i + 1, col[k] + 1, val[k]);
    fclose(f);
    return BRM_OK;
}

}  // extern "C"
