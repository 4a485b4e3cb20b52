// Input validation, per-row intermediate-product counts and row binning.
//
// The per-A-entry product counts and their prefix (scan_alen in scan.hip)
// replace compute_flop (BRMerge-SpGEMM/inc/spgemm.h:31-56); the row counts
// are differences of that prefix.  Binning replaces the reference's
// flop-balanced static_section partition (spgemm.h:247-256, 494-503): instead
// of contiguous ranges per thread, rows are sorted into tiers by product
// count, and each tier runs the kernel shape that suits it.  Binning never
// changes results (each row is merged by the same tree whatever its tier).
//
// Validation replaces the checks the reference's reader performs on load
// (src/CSR.cpp:454-477) and guards the kernels against out-of-range indices.
#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int NT = 256;

// All range checks in one launch: A's and B's row pointers (first 0, last
// nnz, nondecreasing), then A's columns in [0, K) and B's in [0, N).  B is
// skipped when nnzB < 0.  Grid-stride over the concatenated index space.
__global__ __launch_bounds__(NT) void k_validate_all(RowArgs a, int64_t nnzA, int64_t nnzB, int* err) {
    const int64_t n0 = (int64_t)a.M + 1;
    const int64_t n1 = nnzB >= 0 ? (int64_t)a.K + 1 : 0;
    const int64_t n2 = nnzA;
    const int64_t n3 = nnzB > 0 ? nnzB : 0;
    const int64_t total = n0 + n1 + n2 + n3;
    bool bad = false;
    for (int64_t g = (int64_t)blockIdx.x * NT + threadIdx.x; g < total; g += (int64_t)gridDim.x * NT) {
        if (g < n0) {
            const int64_t i = g;
            const int v = a.arpt[i];
            bad |= (i == 0 && v != 0) || (i == a.M && (int64_t)v != nnzA) || (i < a.M && v > a.arpt[i + 1]);
        } else if (g < n0 + n1) {
            const int64_t i = g - n0;
            const int v = a.brpt[i];
            bad |= (i == 0 && v != 0) || (i == a.K && (int64_t)v != nnzB) || (i < a.K && v > a.brpt[i + 1]);
        } else if (g < n0 + n1 + n2) {
            const int v = a.acol[g - n0 - n1];
            bad |= v < 0 || v >= a.K;
        } else {
            const int v = a.bcol[g - n0 - n1 - n2];
            bad |= v < 0 || v >= a.N;
        }
    }
    if (bad) atomicOr(err, (int)DERR_INVALID);
}

// Strictly increasing columns within each row.  Each wavefront checks the
// pairs (k, k + 1) of a fixed chunk of VS_CHUNK entries, 64 at a time, and
// skips a pair where k + 1 starts a row, so a hub row costs no more than
// as many short ones.  The chunk's columns are loaded first, all windows at
// once (col[k + 1] comes from the next lane).  The lanes hold 64 consecutive
// row starts (from the first one after the chunk's first entry, found by a
// 64-way search); a window's row starts are one OR over the lanes, and the
// next 64 are loaded only when a window reaches past the last one held.
constexpr int VS_WIN = 8;
constexpr int VS_CHUNK = VS_WIN * BRM_WAVE;

__global__ void k_validate_sorted(const int* __restrict__ rpt, const int* __restrict__ col, int n, int npairs,
                                  int64_t chunk0, int* err) {
    const int64_t e0 = (chunk0 + (int64_t)blockIdx.x * (NT / BRM_WAVE) + threadIdx.x / BRM_WAVE) * VS_CHUNK;
    if (e0 >= npairs) return;
    if (*err) return;  // row pointers already flagged (checked by the launches before): not followed
    const int lane = lane_id();
    const int x = (int)e0;
    int c[VS_WIN];  // entries 0 .. npairs (= nnz - 1) exist
#pragma unroll
    for (int w = 0; w < VS_WIN; ++w) {
        const int64_t k = e0 + w * BRM_WAVE + lane;
        c[w] = k <= npairs ? col[k] : 0;
    }
    const int cx = e0 + VS_CHUNK <= npairs ? col[e0 + VS_CHUNK] : 0;  // first entry of the next chunk
    // first row start after entry x: the least i in [0, n] with rpt[i] > x
    // (rpt[n] = npairs + 1 > x)
    int lo = 0, hi = n;
    while (lo < hi) {
        const int step = (hi - lo + BRM_WAVE - 1) / BRM_WAVE;
        const int64_t p = lo + (int64_t)lane * step;
        const bool gt = p >= hi || rpt[p] > x;
        const uint64_t m = __ballot(gt);
        if (m == 0ull) {
            lo += (BRM_WAVE - 1) * step + 1;
        } else {
            const int f = __ffsll((long long)m) - 1;
            if (f == 0) break;  // rpt[lo] > x
            const int nlo = lo + (f - 1) * step + 1;
            hi = min(hi, lo + f * step);
            lo = nlo;
        }
    }
    int rb = lo;
    constexpr int64_t NONE = (int64_t)1 << 40;  // past every entry
    int64_t st = rb + lane <= n ? rpt[rb + lane] : NONE;  // start of row rb + lane
    const int eend = (int)min<int64_t>(e0 + VS_CHUNK, npairs);
    bool bad = false;
#pragma unroll
    for (int w = 0; w < VS_WIN; ++w) {  // pairs (k, k + 1), k = kb + lane
        const int64_t kb = e0 + w * BRM_WAVE;
        if (kb >= eend) break;
        // bit j: entry kb + 1 + j starts a row (j = 0 .. 63)
        unsigned blo = 0u, bhi = 0u;
        while (true) {
            const int64_t dd = st - kb - 1;
            const unsigned d = (uint64_t)dd < 64u ? (unsigned)dd : 64u;
            blo |= d < 32u ? 1u << d : 0u;
            bhi |= d - 32u < 32u ? 1u << (d - 32u) : 0u;
            if (__shfl(st, BRM_WAVE - 1, BRM_WAVE) - kb > BRM_WAVE) break;  // the next starts lie past the window
            rb += BRM_WAVE;
            st = rb + lane <= n ? rpt[rb + lane] : NONE;
        }
#pragma unroll
        for (int s = BRM_WAVE / 2; s > 0; s >>= 1) {
            blo |= (unsigned)__shfl_xor((int)blo, s, BRM_WAVE);
            bhi |= (unsigned)__shfl_xor((int)bhi, s, BRM_WAVE);
        }
        const unsigned starts = lane < 32 ? blo >> lane : bhi >> (lane - 32);
        const int first_next = w + 1 < VS_WIN ? __shfl(c[w + 1 < VS_WIN ? w + 1 : w], 0, BRM_WAVE) : cx;
        const int down = __shfl_down(c[w], 1, BRM_WAVE);
        const int nxt = lane == BRM_WAVE - 1 ? first_next : down;
        const int64_t k = kb + lane;
        if (k < eend && !(starts & 1u)) bad |= c[w] >= nxt;
    }
    if (__ballot(bad) != 0ull && lane == 0) atomicOr(err, (int)DERR_INVALID);
}

// Row tiers.  A block takes RS_IT * NT rows; bin slots are counted in LDS and
// claimed with one global atomic per bin and block.
constexpr int RS_IT = 4;

__global__ __launch_bounds__(NT) void k_rowstats(RowArgs a, int med_cap, int med_lmax, int wmed_cap, int force_long,
                                                 int hash_cap, int64_t* __restrict__ iprow,
                                                 int64_t* __restrict__ rowcnt, int* __restrict__ bins,
                                                 int* __restrict__ bin_cnt, const int* __restrict__ err) {
    __shared__ int lcnt[BIN_COUNT];
    __shared__ int gbase[BIN_COUNT];
    const int t = threadIdx.x;
    if (t < BIN_COUNT) lcnt[t] = 0;
    __syncthreads();
    const bool bad = err && *err;  // malformed input: its row pointers may index anywhere
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    int bin[RS_IT], lpos[RS_IT];
#pragma unroll
    for (int it = 0; it < RS_IT; ++it) {
        const int64_t i = ((int64_t)blockIdx.x * RS_IT + it) * NT + t;
        int b_ = -1;
        if (i < a.M && bad) {
            b_ = BIN_EMPTY;
            iprow[i] = 0;
            rowcnt[i] = 0;
        } else if (i < a.M) {
            const int a0 = a.arpt[i], a1 = a.arpt[i + 1];
            const int L = a1 - a0;
            const int64_t ip = a.ipA[a1] - a.ipA[a0];
            iprow[i] = ip;
            if (ip == 0) {
                b_ = BIN_EMPTY;
                rowcnt[i] = 0;
            } else if (force_long) {
                b_ = BIN_LONG;
            } else if (ip <= hash_cap && L <= HASH_CAP) {
                b_ = (ip <= HASH_SMALL && L <= HASH_SMALL) ? BIN_HASH : BIN_HASH2;
            } else if (ip <= TINY_MAX) {
                b_ = BIN_TINY;
            } else if (ip <= SHORT_MAX) {
                b_ = BIN_SHORT;
            } else if (ip <= wmed_cap && L <= WMED_LMAX) {
                b_ = ip <= WMED_SPLIT0 ? BIN_WMED : ip <= WMED_SPLIT ? BIN_WMED2 : ip <= WMED_SPLIT2 ? BIN_WMED3 : BIN_WMED4;
            } else if (ip <= med_cap && L <= med_lmax) {
                b_ = BIN_MEDIUM;
            } else {
                b_ = BIN_LONG;
            }
        }
        bin[it] = b_;
        lpos[it] = 0;
#pragma unroll
        for (int b = 0; b < BIN_COUNT; ++b) {
            const uint64_t m = __ballot(b_ == b);
            if (m == 0ull) continue;
            const int leader = __ffsll((long long)m) - 1;
            int pos = 0;
            if (lane == leader) pos = atomicAdd(&lcnt[b], __popcll(m));
            pos = __shfl(pos, leader, BRM_WAVE);
            if (b_ == b) lpos[it] = pos + __popcll(m & below);
        }
    }
    __syncthreads();
    if (t < BIN_COUNT) gbase[t] = lcnt[t] ? atomicAdd(&bin_cnt[t], lcnt[t]) : 0;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < RS_IT; ++it) {
        const int b = bin[it];
        if (b > 0) bins[(int64_t)(b - 1) * a.M + gbase[b] + lpos[it]] = (int)(((int64_t)blockIdx.x * RS_IT + it) * NT + t);
    }
}

__global__ void k_long_info(const int* __restrict__ rows, int n, RowArgs a, int64_t* ip_out, int* len_out) {
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    const int r = rows[k];
    const int a0 = a.arpt[r], a1 = a.arpt[r + 1];
    ip_out[k] = a.ipA[a1] - a.ipA[a0];
    len_out[k] = a1 - a0;
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_validate(const RowArgs& a, int64_t nnzA, int64_t nnzB, int* err, hipStream_t s) {
    // Row pointers are checked before the sortedness check (a later launch)
    // indexes columns through them.  nnzB < 0: B was validated before.
    const int64_t total = (int64_t)a.M + 1 + (nnzB >= 0 ? (int64_t)a.K + 1 + nnzB : 0) + nnzA;
    {
        ProfScope ps_("k_validate", s);
        k_validate_all<<<(unsigned)std::min<int64_t>(nblk(total, NT), 8192), NT, 0, s>>>(a, nnzA, nnzB, err);
    }
    return hipGetLastError();
}

hipError_t launch_validate_sorted(const int* rpt, const int* col, int n, int64_t nnz, int* err, hipStream_t s) {
    if (n <= 0 || nnz <= 1) return hipSuccess;
    {
        ProfScope ps_("k_validate_sorted", s);
        const int64_t chunks = (nnz - 1 + VS_CHUNK - 1) / VS_CHUNK;  // one wavefront each
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < chunks; b += ch) {
            const int64_t c = std::min<int64_t>(ch, chunks - b);
            k_validate_sorted<<<nblk(c, NT / BRM_WAVE), NT, 0, s>>>(rpt, col, n, (int)(nnz - 1), b, err);
        }
    }
    return hipGetLastError();
}

hipError_t launch_rowstats(const RowArgs& a, int med_cap, int med_lmax, int wmed_cap, int force_long, int hash_cap,
                           int64_t* iprow, int64_t* rowcnt, int* bins, int* bin_cnt, const int* err, hipStream_t s) {
    if (a.M <= 0) return hipSuccess;
    {
        ProfScope ps_("k_rowstats", s);
        k_rowstats<<<nblk(a.M, NT * RS_IT), NT, 0, s>>>(a, med_cap, med_lmax, wmed_cap, force_long,
                                                        std::min(hash_cap, HASH_CAP), iprow, rowcnt, bins, bin_cnt,
                                                        err);
    }
    return hipGetLastError();
}

hipError_t launch_long_info(const int* rows, int n, const RowArgs& a, int64_t* ip_out, int* len_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_long_info", s);
        k_long_info<<<nblk(n, NT), NT, 0, s>>>(rows, n, a, ip_out, len_out);
    }
    return hipGetLastError();
}

}  // namespace brm
