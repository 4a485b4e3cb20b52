// Batched SpGEMM: many independent products C_i = A_i * B_i run as one
// block-diagonal product
//     diag(A_1 .. A_n) * diag(B_1 .. B_n) = diag(C_1 .. C_n).
// A_i's columns are shifted by the first fused row of B_i so that each A row
// selects rows of its own B; B_i's columns are kept, so every fused C row
// carries exactly the columns (and, row by row, the same merge tree and the
// same values) of the corresponding row of C_i.  One set of launches covers
// the whole batch instead of one set per product.
//
// A plan (brm_batch_plan_*) keeps the fused structure, the symbolic result
// and C resident, picks the tier shapes by timing candidates on the batch
// itself, and captures the numeric phase -- value refresh from the caller's
// arrays plus the short and medium tiers -- in one HIP graph that each
// execution replays.
#include <limits.h>
#include <stdio.h>
#include <string.h>

#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "brm_device.h"
#include "pipeline.h"

namespace brm {
namespace {

constexpr int CAT_NT = 256;
constexpr int CAT_ITEMS = 4;

// Device view of one product of the batch.
struct PairDev {
    const int* arpt;
    const int* acol;
    const double* aval;
    const int* brpt;
    const int* bcol;
    const double* bval;
    int64_t nnzA, nnzB;
    int M, K, N;
    int pad;
};

// Prefix tables of the fused layout, each [n+1].
struct Fused {
    const PairDev* pairs;
    int n;
    const int64_t* mbase;  // fused A rows
    const int64_t* kbase;  // fused B rows
    const int64_t* abase;  // fused A entries
    const int64_t* bbase;  // fused B entries
    const int64_t* achunk;  // entry chunks of A: chunks of pair p are achunk[p] .. achunk[p+1]-1
    const int64_t* bchunk;  // entry chunks of B, numbered after all of A's
    int* arpt;
    int* acol;
    double* aval;
    int* brpt;
    int* bcol;
    double* bval;
    int64_t M, K, nnzA, nnzB;
};

// Largest i in [lo, hi] with base[i] <= x (base non-decreasing).
__device__ __forceinline__ int seg_search(const int64_t* __restrict__ base, int lo, int hi, int64_t x) {
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Segment of x for a wavefront whose lanes hold x0 .. x0+63: one search for
// each end of the wavefront's range, then a short walk per lane.
__device__ __forceinline__ int wave_seg(const int64_t* __restrict__ base, int n, int64_t x0, int64_t x) {
    const int plo = seg_search(base, 0, n - 1, x0);
    const int phi = seg_search(base, plo, n - 1, x0 + BRM_WAVE - 1);
    int p = plo;
    while (p < phi && base[p + 1] <= x) ++p;
    return p;
}

// Fused row pointers of A and B (per pair: shifted by the pair's first
// fused entry) and the per-pair checks that the fused arrays cannot show:
// rpt[0] == 0 and rpt[M] == nnz of each operand.
__global__ __launch_bounds__(CAT_NT) void k_cat_rows(Fused f, int* err) {
    const int64_t chunk = ((int64_t)blockIdx.x * CAT_NT + (threadIdx.x & ~(BRM_WAVE - 1))) * CAT_ITEMS;
    const int lane = lane_id();
    const int64_t total = f.M + f.K;
    for (int it = 0; it < CAT_ITEMS; ++it) {
        const int64_t x0 = chunk + (int64_t)it * BRM_WAVE;
        if (x0 >= total) break;
        const int64_t x = x0 + lane;
        if (x < f.M) {
            const int p = x0 + BRM_WAVE <= f.M ? wave_seg(f.mbase, f.n, x0, x) : seg_search(f.mbase, 0, f.n - 1, x);
            const PairDev& P = f.pairs[p];
            const int64_t r = x - f.mbase[p];
            const int v = P.arpt[r];
            f.arpt[x] = (int)(v + f.abase[p]);
            if (r == 0 && v != 0) atomicOr(err, (int)DERR_INVALID);
            if (r == P.M - 1 && (int64_t)P.arpt[P.M] != P.nnzA) atomicOr(err, (int)DERR_INVALID);
            if (v < 0 || (int64_t)v > P.nnzA) atomicOr(err, (int)DERR_INVALID);
        } else if (x < total) {
            const int64_t y = x - f.M;
            const int64_t y0 = x0 - f.M;
            const int p = (y0 >= 0 && x0 + BRM_WAVE <= total) ? wave_seg(f.kbase, f.n, y0, y)
                                                              : seg_search(f.kbase, 0, f.n - 1, y);
            const PairDev& P = f.pairs[p];
            const int64_t r = y - f.kbase[p];
            const int v = P.brpt[r];
            f.brpt[y] = (int)(v + f.bbase[p]);
            if (r == 0 && v != 0) atomicOr(err, (int)DERR_INVALID);
            if (r == P.K - 1 && (int64_t)P.brpt[P.K] != P.nnzB) atomicOr(err, (int)DERR_INVALID);
            if (v < 0 || (int64_t)v > P.nnzB) atomicOr(err, (int)DERR_INVALID);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f.arpt[f.M] = (int)f.nnzA;
        f.brpt[f.K] = (int)f.nnzB;
    }
}

// Fused entries: A columns shifted to the pair's B rows (checked against
// the pair's K), B columns kept (checked against the pair's N); values
// copied when `vals`, indices written when `cols`.  Entries go in chunks of
// CAT_CHUNK, each within one pair: workgroup b finds its pair by a search
// over the chunk table (the same for the whole workgroup) and copies its
// chunk with coalesced loads and stores, CAT_EPT entries per thread in
// flight.
constexpr int CAT_EPT = 16;
constexpr int64_t CAT_CHUNK = (int64_t)CAT_NT * CAT_EPT;

__global__ __launch_bounds__(CAT_NT) void k_cat_entries(Fused f, int cols, int vals, int* err) {
    const int64_t b = blockIdx.x;
    const int64_t na = f.achunk[f.n];
    const bool isA = b < na;
    const int64_t* chunk = isA ? f.achunk : f.bchunk;
    const int64_t cb = isA ? b : b - na;  // chunk number within A's or B's
    const int p = seg_search(chunk, 0, f.n - 1, cb);
    const PairDev& P = f.pairs[p];
    const int64_t j0 = (cb - chunk[p]) * CAT_CHUNK;
    const int64_t jn = min((int64_t)CAT_CHUNK, (isA ? P.nnzA : P.nnzB) - j0);
    const int* sc = isA ? P.acol : P.bcol;
    const double* sv = isA ? P.aval : P.bval;
    int* dc = (isA ? f.acol + f.abase[p] : f.bcol + f.bbase[p]) + j0;
    double* dv = (isA ? f.aval + f.abase[p] : f.bval + f.bbase[p]) + j0;
    const int lim = isA ? P.K : P.N;
    const int shift = isA ? (int)f.kbase[p] : 0;
    bool bad = false;
    if (cols) {
        int c[CAT_EPT];
#pragma unroll
        for (int i = 0; i < CAT_EPT; ++i) {
            const int64_t j = (int64_t)i * CAT_NT + threadIdx.x;
            c[i] = j < jn ? sc[j0 + j] : 0;
        }
#pragma unroll
        for (int i = 0; i < CAT_EPT; ++i) {
            const int64_t j = (int64_t)i * CAT_NT + threadIdx.x;
            if (j < jn) {
                bad |= c[i] < 0 || c[i] >= lim;
                dc[j] = min(max(c[i], 0), max(lim - 1, 0)) + shift;
            }
        }
    }
    if (vals) {
        double v[CAT_EPT];
#pragma unroll
        for (int i = 0; i < CAT_EPT; ++i) {
            const int64_t j = (int64_t)i * CAT_NT + threadIdx.x;
            v[i] = j < jn ? sv[j0 + j] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < CAT_EPT; ++i) {
            const int64_t j = (int64_t)i * CAT_NT + threadIdx.x;
            if (j < jn) dv[j] = v[i];
        }
    }
    if (__ballot(bad) != 0ull && lane_id() == 0) atomicOr(err, (int)DERR_INVALID);
}

// Per-product row pointers of a fused result: product p owns fused rows
// mbase[p] .. mbase[p+1]-1 and entries ebase[p] .. ebase[p] + M_p of
// rpt_out, which restart at 0; base_out[p] is the product's first entry in
// the fused column and value arrays.
__global__ __launch_bounds__(CAT_NT) void k_split_rpt(const int64_t* __restrict__ crpt,
                                                      const int64_t* __restrict__ mbase,
                                                      const int64_t* __restrict__ ebase, int n, int64_t total,
                                                      int64_t* __restrict__ rpt_out, int64_t* __restrict__ base_out) {
    const int64_t y = (int64_t)blockIdx.x * CAT_NT + threadIdx.x;
    if (y >= total) return;
    const int p = seg_search(ebase, 0, n - 1, y);
    const int64_t i = y - ebase[p];
    const int64_t b = crpt[mbase[p]];
    rpt_out[y] = crpt[mbase[p] + i] - b;
    if (i == 0) base_out[p] = b;
}

unsigned cat_blocks(int64_t n) {
    const int64_t per = (int64_t)CAT_NT * CAT_ITEMS;
    return (unsigned)((n + per - 1) / per);
}

// ---------------------------------------------------------------- host side
// The fused operands of one batch (device), built from the caller's pairs.
struct FusedBatch {
    Scratch sc;
    Fused f{};
    int Nmax = 0;
    int64_t nchunks = 0;  // workgroups of k_cat_entries
    std::vector<int64_t> mbase, kbase, abase, bbase;
};

int check_pairs(const brm_pair* pairs, int n) {
    if (n < 0 || (n > 0 && !pairs)) {
        set_err("invalid pair array");
        return BRM_E_INVALID;
    }
    for (int i = 0; i < n; ++i) {
        const brm_pair& q = pairs[i];
        if (q.M < 0 || q.K < 0 || q.N < 0 || q.nnzA < 0 || q.nnzB < 0) {
            set_err("pair %d: negative dimension", i);
            return BRM_E_INVALID;
        }
        if ((q.M > 0 && !q.arpt) || (q.K > 0 && !q.brpt) || (q.nnzA > 0 && (!q.acol || !q.aval)) ||
            (q.nnzB > 0 && (!q.bcol || !q.bval))) {
            set_err("pair %d: null pointer", i);
            return BRM_E_INVALID;
        }
        if ((q.M == 0 && q.nnzA != 0) || (q.K == 0 && q.nnzB != 0)) {
            set_err("pair %d: entries without rows", i);
            return BRM_E_INVALID;
        }
    }
    return BRM_OK;
}

int build_fused(const brm_pair* pairs, int n, hipStream_t s, FusedBatch& fb, int* derr, const int64_t* a_col_off) {
    BRMC(check_pairs(pairs, n));
    fb.mbase.assign(n + 1, 0);
    fb.kbase.assign(n + 1, 0);
    fb.abase.assign(n + 1, 0);
    fb.bbase.assign(n + 1, 0);
    std::vector<PairDev> pd(n > 0 ? n : 1);
    for (int i = 0; i < n; ++i) {
        const brm_pair& q = pairs[i];
        fb.mbase[i + 1] = fb.mbase[i] + q.M;
        fb.kbase[i + 1] = fb.kbase[i] + q.K;
        fb.abase[i + 1] = fb.abase[i] + q.nnzA;
        fb.bbase[i + 1] = fb.bbase[i] + q.nnzB;
        fb.Nmax = std::max(fb.Nmax, q.N);
        pd[i] = PairDev{q.arpt, q.acol, q.aval, q.brpt, q.bcol, q.bval, q.nnzA, q.nnzB, q.M, q.K, q.N, 0};
    }
    Fused& f = fb.f;
    f.n = n;
    f.M = fb.mbase[n];
    f.K = fb.kbase[n];
    f.nnzA = fb.abase[n];
    f.nnzB = fb.bbase[n];
    if (a_col_off) {
        // B operands are fused elsewhere: only A is assembled, with the
        // caller's per-pair column offsets into that fused B
        for (int i = 0; i < n; ++i) fb.kbase[i] = a_col_off[i];
        f.K = 0;
        f.nnzB = 0;
    }
    if (f.M > INT32_MAX - 1 || f.K > INT32_MAX - 1 || f.nnzA > INT32_MAX || f.nnzB > INT32_MAX) {
        set_err("fused batch exceeds int32 rows or entries (%lld rows, %lld + %lld entries)", (long long)f.M,
                (long long)f.nnzA, (long long)f.nnzB);
        return BRM_E_INVALID;
    }
    PairDev* pairs_d;
    int64_t* tab_d;
    Scratch& sc = fb.sc;
    BRMC(sc.alloc(&pairs_d, n, "batch pairs"));
    BRMC(sc.alloc(&tab_d, 6 * ((int64_t)n + 1), "batch tables"));
    BRMC(sc.alloc(&f.arpt, f.M + 1, "fused A rpt"));
    BRMC(sc.alloc(&f.acol, f.nnzA, "fused A col"));
    BRMC(sc.alloc(&f.aval, f.nnzA, "fused A val"));
    BRMC(sc.alloc(&f.brpt, f.K + 1, "fused B rpt"));
    BRMC(sc.alloc(&f.bcol, f.nnzB, "fused B col"));
    BRMC(sc.alloc(&f.bval, f.nnzB, "fused B val"));
    // entry chunks: those of A first, then those of B
    std::vector<int64_t> achunk(n + 1, 0), bchunk(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        achunk[i + 1] = achunk[i] + (pairs[i].nnzA + CAT_CHUNK - 1) / CAT_CHUNK;
        bchunk[i + 1] = bchunk[i] + (a_col_off ? 0 : (pairs[i].nnzB + CAT_CHUNK - 1) / CAT_CHUNK);
    }
    fb.nchunks = achunk[n] + bchunk[n];
    std::vector<int64_t> tab;
    tab.reserve(6 * ((size_t)n + 1));
    for (auto* v : {&fb.mbase, &fb.kbase, &fb.abase, &fb.bbase, &achunk, &bchunk})
        tab.insert(tab.end(), v->begin(), v->end());
    if (n > 0) HIPC(hipMemcpyAsync(pairs_d, pd.data(), sizeof(PairDev) * n, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(tab_d, tab.data(), sizeof(int64_t) * tab.size(), hipMemcpyHostToDevice, s));
    f.pairs = pairs_d;
    f.mbase = tab_d;
    f.kbase = tab_d + (n + 1);
    f.abase = tab_d + 2 * ((int64_t)n + 1);
    f.bbase = tab_d + 3 * ((int64_t)n + 1);
    f.achunk = tab_d + 4 * ((int64_t)n + 1);
    f.bchunk = tab_d + 5 * ((int64_t)n + 1);
    HIPC(launch_fill_i32(derr, 1, 0, s));
    if (n > 0) {
        {
            ProfScope ps_("k_cat_rows", s);
            k_cat_rows<<<cat_blocks(f.M + f.K), CAT_NT, 0, s>>>(f, derr);
        }
        HIPC(hipGetLastError());
        if (fb.nchunks > 0) {
            ProfScope ps_("k_cat_entries", s);
            k_cat_entries<<<(unsigned)fb.nchunks, CAT_NT, 0, s>>>(f, 1, 1, derr);
        }
        HIPC(hipGetLastError());
    } else {
        HIPC(launch_fill_i32(f.arpt, 1, 0, s));
        HIPC(launch_fill_i32(f.brpt, 1, 0, s));
    }
    return check_dev_err(derr, s, "batch assembly");
}

DevIn fused_in(const FusedBatch& fb) {
    const Fused& f = fb.f;
    return DevIn{f.arpt, f.acol, f.aval, f.nnzA, f.brpt, f.bcol, f.bval, f.nnzB, (int)f.M, (int)f.K, fb.Nmax};
}

hipError_t refresh_values(const FusedBatch& fb, int* derr, hipStream_t s) {
    const Fused& f = fb.f;
    const int64_t nchunks = fb.nchunks;
    if (f.n == 0 || nchunks == 0) return hipSuccess;
    {
        ProfScope ps_("k_cat_entries", s);
        k_cat_entries<<<(unsigned)nchunks, CAT_NT, 0, s>>>(f, 0, 1, derr);
    }
    return hipGetLastError();
}

// Tier-shape candidates tried by the plan's tuner: (LDS shape, medium
// capacity, wavefront-tier capacity).
struct TuneCand {
    int lds_cfg;
    int med_cap;
    int wmed_cap;
};
const TuneCand kCands[] = {{0, 4096, 0},   {2, 2048, 0},   {0, 4096, 256}, {0, 4096, 512},
                           {0, 4096, 1024}, {2, 2048, 512}, {2, 2048, 1024}};

// ---- tuned-shape cache: a JSON file of tier shapes keyed by a batch's
// shape statistics (log2 buckets of rows, products, columns and the row
// count of every bin), so that a plan over a batch of the same shape skips
// the tuner.  Written one entry per line (read in any JSON layout):
//   {"version": 1, "entries": [
//     {"key": "...", "lds_cfg": 0, "med_cap": 4096, "wmed_cap": 1024, "ms": 250.9},
//     ...
//   ]}
// The file is rewritten whole: a temporary file, then rename.
std::mutex g_tune_mu;
std::string g_tune_path;
bool g_tune_path_set = false;

std::string tune_path() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_path_set) {
        const char* e = getenv("BRM_TUNE_CACHE");
        g_tune_path = e ? e : "";
        g_tune_path_set = true;
    }
    return g_tune_path;
}

int lg2_bucket(int64_t x) {
    int b = 0;
    while (x > 1) {
        x >>= 1;
        ++b;
    }
    return b;
}

std::string tune_key(const Pattern& pat) {
    std::string k = "rows" + std::to_string(lg2_bucket(pat.a.M)) + "_ip" + std::to_string(lg2_bucket(pat.total_ip)) +
                    "_n" + std::to_string(lg2_bucket(pat.a.N)) + "_bins";
    for (int b = 0; b < BIN_COUNT; ++b) k += (b ? "." : "") + std::to_string(lg2_bucket((int64_t)pat.bc[b] + 1));
    return k;
}

struct TuneEntry {
    std::string key;
    int lds_cfg, med_cap, wmed_cap;
    double ms;
};

// Value of "name" (a number) in one entry object's text, or false.
bool json_num(const std::string& obj, const char* name, double* v) {
    const std::string pat = std::string("\"") + name + "\"";
    const size_t at = obj.find(pat);
    if (at == std::string::npos) return false;
    size_t i = at + pat.size();
    while (i < obj.size() && (obj[i] == ' ' || obj[i] == ':' || obj[i] == '\t' || obj[i] == '\n')) ++i;
    const char* p = obj.c_str() + i;
    char* end = nullptr;
    *v = strtod(p, &end);
    return end != p;
}

// Every entry object of the file: the text from each "key" to the next '}'.
std::vector<TuneEntry> tune_load(const std::string& path) {
    std::vector<TuneEntry> out;
    std::ifstream in(path);
    if (!in) return out;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    for (size_t at = text.find("\"key\""); at != std::string::npos; at = text.find("\"key\"", at + 5)) {
        const size_t close = text.find('}', at);
        if (close == std::string::npos) break;
        const std::string obj = text.substr(at, close - at);
        size_t i = 5;
        while (i < obj.size() && (obj[i] == ' ' || obj[i] == ':' || obj[i] == '\t' || obj[i] == '\n')) ++i;
        if (i >= obj.size() || obj[i] != '"') continue;
        const size_t ke = obj.find('"', i + 1);
        if (ke == std::string::npos) continue;
        TuneEntry e{obj.substr(i + 1, ke - i - 1), 0, 0, 0, 0.0};
        double a, b, c, ms = 0;
        if (!json_num(obj, "lds_cfg", &a) || !json_num(obj, "med_cap", &b) || !json_num(obj, "wmed_cap", &c)) continue;
        (void)json_num(obj, "ms", &ms);
        e.lds_cfg = (int)a;
        e.med_cap = (int)b;
        e.wmed_cap = (int)c;
        e.ms = ms;
        out.push_back(e);
    }
    return out;
}

// The candidate a cached entry names, or -1 (absent, or not a candidate).
int tune_lookup(const std::string& path, const std::string& key) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const TuneEntry& e : tune_load(path)) {
        if (e.key != key) continue;
        for (int c = 0; c < (int)(sizeof(kCands) / sizeof(kCands[0])); ++c)
            if (kCands[c].lds_cfg == e.lds_cfg && kCands[c].med_cap == e.med_cap && kCands[c].wmed_cap == e.wmed_cap)
                return c;
    }
    return -1;
}

int tune_store(const std::string& path, const std::string& key, int cand, double ms) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::vector<TuneEntry> es = tune_load(path);
    const TuneCand& c = kCands[cand];
    bool found = false;
    for (TuneEntry& e : es) {
        if (e.key == key) {
            e = TuneEntry{key, c.lds_cfg, c.med_cap, c.wmed_cap, ms};
            found = true;
        }
    }
    if (!found) es.push_back(TuneEntry{key, c.lds_cfg, c.med_cap, c.wmed_cap, ms});
    const std::string tmp = path + ".tmp";
    {
        std::ofstream o(tmp, std::ios::trunc);
        if (!o) {
            set_err("tune cache: cannot write %s", tmp.c_str());
            return BRM_E_INVALID;
        }
        o << "{\"version\": 1, \"entries\": [\n";
        for (size_t i = 0; i < es.size(); ++i) {
            char item[512];
            snprintf(item, sizeof(item),
                     "  {\"key\": \"%s\", \"lds_cfg\": %d, \"med_cap\": %d, \"wmed_cap\": %d, \"ms\": %.6f}%s\n",
                     es[i].key.c_str(), es[i].lds_cfg, es[i].med_cap, es[i].wmed_cap, es[i].ms,
                     i + 1 < es.size() ? "," : "");
            o << item;
        }
        o << "]}\n";
        if (!o) {
            set_err("tune cache: write to %s failed", tmp.c_str());
            return BRM_E_INVALID;
        }
    }
    if (rename(tmp.c_str(), path.c_str()) != 0) {
        set_err("tune cache: cannot rename %s to %s", tmp.c_str(), path.c_str());
        return BRM_E_INVALID;
    }
    return BRM_OK;
}

}  // namespace

int fuse_pairs(const brm_pair* pairs, int n, const int64_t* a_col_off, hipStream_t s, FusedOut* out) {
    FusedBatch fb;
    int* derr;
    BRMC(fb.sc.alloc(&derr, 1, "error flag"));
    BRMC(build_fused(pairs, n, s, fb, derr, a_col_off));
    out->in = fused_in(fb);
    out->mbase = fb.mbase;
    out->sc.bufs.swap(fb.sc.bufs);
    return BRM_OK;
}

hipError_t launch_split_rpt(const int64_t* crpt, const int64_t* mbase_d, const int64_t* ebase_d, int n,
                            int64_t total, int64_t* rpt_out, int64_t* base_out, hipStream_t s) {
    if (n <= 0 || total <= 0) return hipSuccess;
    {
        ProfScope ps_("k_split_rpt", s);
        k_split_rpt<<<(unsigned)((total + CAT_NT - 1) / CAT_NT), CAT_NT, 0, s>>>(crpt, mbase_d, ebase_d, n, total,
                                                                                  rpt_out, base_out);
    }
    return hipGetLastError();
}

}  // namespace brm

using namespace brm;

struct brm_batch_plan {
    FusedBatch fb;
    Pattern pat;
    int* ccol = nullptr;
    double* cval = nullptr;
    std::vector<int64_t> row_off;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<int, double>> tuned;  // candidate index, best ms
    int chosen = -1;
    std::string tune_key;
    const char* tune_source = "none";  // "timed", "cache" or "none"
    double analyse_s = 0;
    ~brm_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

namespace {

int batched_run(const brm_pair* pairs, int n, hipStream_t s, DevOut* out, int64_t* row_off, double* ph) {
    BRMC(check_pairs(pairs, n));
    BRMC(use_device());
    FusedBatch fb;
    int* derr;
    BRMC(fb.sc.alloc(&derr, 1, "error flag"));
    BRMC(build_fused(pairs, n, s, fb, derr, nullptr));
    BRMC(run_spgemm(fused_in(fb), false, INT64_MAX, s, out, ph));
    if (row_off)
        for (int i = 0; i <= n; ++i) row_off[i] = fb.mbase[i];
    return BRM_OK;
}

double time_numeric(brm_batch_plan* p, hipStream_t s, int reps) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1;
    }
    double best = -1;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(a, s);
        if (pattern_numeric_tiers(p->pat, p->ccol, p->cval, s) != BRM_OK) break;
        if (pattern_numeric_long(p->pat, p->ccol, p->cval, s) != BRM_OK) break;
        (void)hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (best < 0 || ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

int plan_build(brm_batch_plan* p, const brm_pair* pairs, int n, int autotune, hipStream_t s) {
    const double t0 = now_s();
    BRMC(check_pairs(pairs, n));
    BRMC(use_device());
    int* derr;
    BRMC(p->fb.sc.alloc(&derr, 1, "error flag"));
    BRMC(build_fused(pairs, n, s, p->fb, derr, nullptr));
    p->row_off = p->fb.mbase;
    Pattern& pat = p->pat;
    BRMC(pattern_prepare(fused_in(p->fb), s, nullptr, pat));
    BRMC(pattern_symbolic(pat, s, nullptr));
    BRMC(pat.sc.alloc(&p->ccol, pat.nnzC, "C col"));
    BRMC(pat.sc.alloc(&p->cval, pat.nnzC, "C val"));
    if (!autotune) {
        // one eager pass: validates the numeric phase outside the graph and
        // initialises every kernel's attributes before capture
        if (time_numeric(p, s, 1) < 0) {
            set_err("batch plan: numeric pass failed");
            return BRM_E_DEVICE;
        }
        BRMC(check_dev_err(pat.derr, s, "batch plan"));
    } else {
        const std::string path = tune_path();
        p->tune_key = tune_key(pat);
        p->chosen = path.empty() ? -1 : tune_lookup(path, p->tune_key);
        double best = -1;
        const int nc = p->chosen >= 0 ? 0 : (int)(sizeof(kCands) / sizeof(kCands[0]));
        p->tune_source = p->chosen >= 0 ? "cache" : "timed";
        for (int c = 0; c < nc; ++c) {
            BRMC(pattern_bin(pat, kCands[c].lds_cfg, kCands[c].med_cap, kCands[c].wmed_cap, 0, s));
            const double ms = time_numeric(p, s, 2);
            BRMC(check_dev_err(pat.derr, s, "batch tuning"));
            if (ms < 0) {
                set_err("batch tuning: timing failed");
                return BRM_E_DEVICE;
            }
            p->tuned.push_back({c, ms});
            if (best < 0 || ms < best) {
                best = ms;
                p->chosen = c;
            }
        }
        if (nc > 0 && !path.empty()) BRMC(tune_store(path, p->tune_key, p->chosen, best));
        const TuneCand& best_c = kCands[p->chosen];
        BRMC(pattern_bin(pat, best_c.lds_cfg, best_c.med_cap, best_c.wmed_cap, 0, s));
        if (nc == 0) {
            // a cached shape: one eager pass, as without the tuner
            if (time_numeric(p, s, 1) < 0) {
                set_err("batch plan: numeric pass failed");
                return BRM_E_DEVICE;
            }
            BRMC(check_dev_err(pat.derr, s, "batch plan"));
        }
    }
    // The graph: value refresh from the caller's arrays, then the short and
    // medium tiers.  Kernels only (no memsets, no host round trips).
    HIPC(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
    Params& prm = params();
    const int prof = prm.profile;
    prm.profile = 0;  // no event records inside the captured sequence
    hipError_t e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal);
    int st = hip_status(e, "hipStreamBeginCapture");
    if (st == BRM_OK) {
        hipError_t e1 = refresh_values(p->fb, pat.derr, p->cap);
        int st1 = e1 == hipSuccess ? pattern_numeric_tiers(pat, p->ccol, p->cval, p->cap) : hip_status(e1, "refresh");
        hipError_t e2 = hipStreamEndCapture(p->cap, &p->graph);
        st = st1 != BRM_OK ? st1 : hip_status(e2, "hipStreamEndCapture");
    }
    prm.profile = prof;
    BRMC(st);
    HIPC(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    HIPC(hipStreamSynchronize(s));
    p->analyse_s = now_s() - t0;
    return BRM_OK;
}

}  // namespace

extern "C" {

int brm_spgemm_batched_dev(const brm_pair* pairs, int npairs, int64_t** crpt, int** ccol, double** cval,
                           int64_t* cnnz, int64_t* row_off, brm_precise_timing* timing, void* stream) {
    if (!crpt || !ccol || !cval || !cnnz) {
        set_err("null output pointer");
        return BRM_E_INVALID;
    }
    const double t0 = now_s();
    double ph[P_N] = {0};
    DevOut out{};
    int st = batched_run(pairs, npairs, (hipStream_t)stream, &out, row_off, ph);
    if (st == BRM_OK) {
        *crpt = out.crpt;
        *ccol = out.ccol;
        *cval = out.cval;
        *cnnz = out.nnz;
        if (timing) {
            memset(timing, 0, sizeof(*timing));
            timing->pre_allocate = ph[P_PRE];
            timing->compute_flop = ph[P_FLOP];
            timing->load_balance = ph[P_LB];
            timing->symbolic = ph[P_SYM];
            timing->prefix_sum_nnz = ph[P_PNNZ];
            timing->allocate_C = ph[P_ALLOC];
            timing->compute = ph[P_COMPUTE];
            timing->cleanup = ph[P_CLEAN];
            timing->total = now_s() - t0;
        }
    }
    dev_flush_deferred();
    return st;
}

int brm_batch_plan_create(const brm_pair* pairs, int npairs, int autotune, brm_batch_plan** plan, void* stream) {
    if (!plan) {
        set_err("null plan pointer");
        return BRM_E_INVALID;
    }
    *plan = nullptr;
    brm_batch_plan* p = new (std::nothrow) brm_batch_plan();
    if (!p) {
        set_err("host allocation failed");
        return BRM_E_NOMEM;
    }
    const int st = plan_build(p, pairs, npairs, autotune, (hipStream_t)stream);
    if (st != BRM_OK) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        delete p;
        dev_flush_deferred();
        return st;
    }
    *plan = p;
    dev_flush_deferred();
    return BRM_OK;
}

int brm_batch_plan_execute(brm_batch_plan* p, void* stream) {
    if (!p) {
        set_err("null plan");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    hipStream_t s = (hipStream_t)stream;
    HIPC(hipGraphLaunch(p->exec, s));
    BRMC(pattern_numeric_long(p->pat, p->ccol, p->cval, s));
    HIPC(hipStreamSynchronize(s));
    const int st = check_dev_err(p->pat.derr, s, "batch execute");
    dev_flush_deferred();
    return st;
}

int brm_batch_plan_result(const brm_batch_plan* p, const int64_t** crpt, const int** ccol, const double** cval,
                          int64_t* cnnz, const int64_t** row_off) {
    if (!p || !crpt || !ccol || !cval || !cnnz || !row_off) {
        set_err("null argument");
        return BRM_E_INVALID;
    }
    *crpt = p->pat.crpt;
    *ccol = p->ccol;
    *cval = p->cval;
    *cnnz = p->pat.nnzC;
    *row_off = p->row_off.data();
    return BRM_OK;
}

int brm_batch_plan_info(const brm_batch_plan* p, char* buf, uint64_t cap) {
    if (!p || !buf) return BRM_E_INVALID;
    std::string s = "{";
    char item[256];
    snprintf(item, sizeof(item),
             "\"pairs\": %d, \"rows\": %lld, \"total_ip\": %lld, \"nnz_c\": %lld, \"rows_by_tier\": [%d, %d, %d, %d], "
             "\"analyse_s\": %.6f, \"lds_cfg\": %d, \"med_cap\": %d, \"wmed_cap\": %d, \"tuning_ms\": [",
             p->fb.f.n, (long long)p->fb.f.M, (long long)p->pat.total_ip, (long long)p->pat.nnzC, p->pat.bc[0],
             p->pat.bc[1] + p->pat.bc[BIN_TINY], p->pat.bc[2] + p->pat.bc[BIN_WMED] + p->pat.bc[BIN_WMED2] + p->pat.bc[BIN_WMED3] + p->pat.bc[BIN_WMED4], p->pat.bc[3], p->analyse_s, p->pat.lds_cfg,
             p->pat.med_cap, p->pat.wmed_cap);
    s += item;
    for (size_t i = 0; i < p->tuned.size(); ++i) {
        const TuneCand& c = kCands[p->tuned[i].first];
        snprintf(item, sizeof(item), "%s[%d, %d, %d, %.6f]", i ? ", " : "", c.lds_cfg, c.med_cap, c.wmed_cap,
                 p->tuned[i].second);
        s += item;
    }
    s += "], \"tune_key\": \"" + p->tune_key + "\", \"tune_source\": \"" + p->tune_source + "\"}";
    if (s.size() + 1 > cap) return BRM_E_INVALID;
    memcpy(buf, s.c_str(), s.size() + 1);
    return BRM_OK;
}

int brm_set_tune_cache(const char* path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_path = path ? path : "";
    g_tune_path_set = true;
    return BRM_OK;
}

void brm_batch_plan_destroy(brm_batch_plan* p) {
    if (!p) return;
    (void)use_device();
    (void)hipDeviceSynchronize();
    delete p;
    dev_flush_deferred();
}

}  // extern "C"
