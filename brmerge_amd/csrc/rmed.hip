// Register tier for small medium rows: one wavefront per row, the row's
// products held in registers (R per lane, ip <= 64 R <= 2048, at most 64
// lists; R = 32 only with bitmap ranks).
//
//   1. every lane l < L holds list l's product offset, B row start and A
//      value; the non-empty lists mark their first product in an LDS bitmap
//      and describe themselves in an LDS table, so each product finds its
//      list by a popcount rank and loads its B entry; value = a * b (rounded,
//      BRMerge-SpGEMM/inc/spgemm.h:621).
//   Sort variant (R <= 16): keys = (column, list position, index), 32-bit
//   when the column fits beside 6 bits of list position and log2(64 R) bits
//   of index (half the shuffles and compares), else 64-bit;
//   2. a bitonic sort over the 64 R keys (cross-lane steps with shuffles,
//      intra-lane steps between registers) orders the products by column,
//      and within a column by list position;
//   3. the sorted keys and the values go to LDS; the first product of every
//      column run sums its run as the aligned binary tree over list
//      positions does (siblings added left + right, a node with one present
//      child passes it through, spgemm.h:639-707) -- a single pass that
//      keeps one pending left subtree per height;
//   4. run heads are compacted with ballots and written to C (Precise), or
//      to C-bar with the row length (Upper).
//   Bitmap variants (B of at most RB2_SPAN columns, R >= 4): the columns set
//   bits of an LDS bitmap -- flat over the row's column span in chunks of
//   2^15 columns (B of at most RB_SPAN columns), or in two levels (level 1
//   numbers the 64-column words in use, level 2 holds those words) -- and
//   popcount ranks give each product its output position; products (still
//   in registers) are bucketed by position into the ranks' LDS, and each
//   bucket is folded in list-position order -- no sort.
// The symbolic count of these rows (k_rmed_count) uses the same list lookup
// and bitmaps; B of more than RB2_SPAN columns takes the wavefront hash
// kernel of lds.hip.
#include <limits.h>

#include <algorithm>
#include <type_traits>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int RM_LOGL = 6;  // list positions < 64: six tree levels

__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t x, int m) {
    const int lo = __shfl_xor((int)(uint32_t)x, m, BRM_WAVE);
    const int hi = __shfl_xor((int)(uint32_t)(x >> 32), m, BRM_WAVE);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t x, int m) {
    return (uint32_t)__shfl_xor((int)x, m, BRM_WAVE);
}

// Key layout: column | list position (LP_BITS) | product index (J_BITS).
template <int R, typename KT>
struct RmKey {
    static constexpr int J_BITS = sizeof(KT) == 8 ? 16 : (R == 2 ? 7 : R == 4 ? 8 : R == 8 ? 9 : 10);
    static constexpr int LP_BITS = sizeof(KT) == 8 ? 16 : 6;
    static constexpr int COL_SHIFT = J_BITS + LP_BITS;
    // Largest column a key holds.  A valid key is all ones only as the
    // last product (j = 64 R - 1) of a row of 64 R products, which has no
    // all-ones filler keys.
    static constexpr int64_t COL_MAX = sizeof(KT) == 8 ? INT32_MAX : ((int64_t)1 << (32 - COL_SHIFT)) - 1;
    __device__ static KT make(int col, int lp, int j) {
        return ((KT)(uint32_t)col << COL_SHIFT) | ((KT)lp << J_BITS) | (KT)j;
    }
    __device__ static uint32_t col(KT k) { return (uint32_t)(k >> COL_SHIFT); }
    __device__ static int lp(KT k) { return (int)((k >> J_BITS) & (((KT)1 << LP_BITS) - 1)); }
    __device__ static int j(KT k) { return (int)(k & (((KT)1 << J_BITS) - 1)); }
};

template <int R, typename KT>
struct RmSmem {
    double val[64 * R];  // first: the bitmap path's val is the same array
    KT key[64 * R];
};

// Stage 1: the row's non-empty lists, k-th at ent[k] (first product, list
// position, B row start, A value); bit j of start is set when a non-empty
// list begins at product j.  Overlays the start of each variant's LDS.
struct RmListEnt {
    int off_lp;  // first product | list position << 16
    int bst;
    double sc;
};
struct RmLists {
    unsigned long long start[32];
    RmListEnt ent[64];
};
static_assert(sizeof(RmLists) <= sizeof(RmSmem<2, uint32_t>), "list table fits the smallest variant");

// Bitmap variants: the row's columns set bits of an LDS bitmap, popcount
// ranks give every product its output position, products are bucketed by
// position and each bucket is summed in list-position order -- O(ip) work
// instead of a sort.  Phase A (ranks) and phase B (buckets) share the LDS;
// the products stay in registers until phase B places them.
//   flat (B with at most RB_SPAN columns): one bitmap over the row's column
//     span, in chunks of RB_WORDS words (two at most);
//   two-level (B with at most RB2_SPAN columns): level 1 has one bit per
//     64-column word of the span, and its popcount ranks number the words
//     in use; level 2 holds those words only (at most one per product).
constexpr int RB_SPAN = 1 << 16;   // flat: columns of B
constexpr int RB_WORDS = 512;      // flat: bitmap words of one chunk of the row's column span
constexpr int RB2_SPAN = 1 << 20;  // two-level: columns of B
constexpr int RB2_L1 = RB2_SPAN / 64 / 64;  // level-1 words
struct RbRanks {
    unsigned long long bm[RB_WORDS];
    unsigned short wrank[RB_WORDS];
};
template <int R>
struct Rb2Ranks {
    unsigned long long l1[RB2_L1];
    unsigned short l1rank[RB2_L1];
    unsigned long long l2[64 * R];
    unsigned short l2rank[64 * R];
};
template <int R, bool TWO>
struct RbBuckets {
    double val[64 * R];    // contributions, bucket by bucket
    unsigned cnt[32 * R];  // 16-bit bucket counters, two per word (bucket 2i low, 2i + 1 high)
    // bucket column (two-level) or bucket column minus the span's first column (flat)
    typename std::conditional<TWO, int, unsigned short>::type oc[64 * R];
    unsigned char lp[64 * R];  // list position of each contribution
};
template <int R, bool TWO>
union RbSmem {
    typename std::conditional<TWO, Rb2Ranks<R>, RbRanks>::type a;
    RbBuckets<R, TWO> b;
    RmLists l;
};
static_assert(64 * 32 <= 0xffff, "bucket offsets in 16 bits");

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Inclusive scan over the wavefront with DPP moves (Hillis-Steele within
// rows of 16 lanes, then row broadcasts): no LDS round trips.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_add(int x) {
    return x + __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ int wave_incl_scan_dpp(int x) {
    x = dpp_add<0x111, 0xF>(x);  // row_shr:1
    x = dpp_add<0x112, 0xF>(x);  // row_shr:2
    x = dpp_add<0x114, 0xF>(x);  // row_shr:4
    x = dpp_add<0x118, 0xF>(x);  // row_shr:8
    x = dpp_add<0x142, 0xA>(x);  // row_bcast15 into rows 1 and 3
    x = dpp_add<0x143, 0xC>(x);  // row_bcast31 into rows 2 and 3
    return x;
}

// Aligned-tree fold over leaves visited in increasing list position: keeps
// one pending left subtree per height (siblings added left + right, a node
// with one present child passes it through).
struct TreeFold {
    double lv[RM_LOGL];
    double cur;
    unsigned pend = 0u;
    int prev;
    __device__ TreeFold(int l, double v) : cur(v), prev(l) {}
    __device__ void push(int l, double x) {
        const int d = 31 - __clz(prev ^ l);
        double v = cur;
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if (h < d && ((pend >> h) & 1u)) v = fadd(lv[h], v);
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if (h == d) lv[h] = v;
        pend = (pend & ~((1u << d) - 1u)) | (1u << d);
        cur = x;
        prev = l;
    }
    __device__ double finish() const {
        double v = cur;
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if ((pend >> h) & 1u) v = fadd(lv[h], v);
        return v;
    }
};

// Sum of one bucket (m >= 3 contributions from distinct lists, list
// positions lp[0..m), values val[0..m)).  Contributions placed by one
// register step of the wavefront may be out of list order; one pass folds
// a bucket that is in order, others visit their leaves by a search.
__device__ double rb_bucket_sum(const unsigned char* lp, const double* val, int m) {
    {
        TreeFold f(lp[0], val[0]);
        int t = 1;
        for (; t < m; ++t) {
            const int l = lp[t];
            if (l < f.prev) break;
            f.push(l, val[t]);
        }
        if (t == m) return f.finish();
    }
    unsigned long long mask = 0ull;
    for (int t = 0; t < m; ++t) mask |= 1ull << lp[t];
    auto leaf = [&](int l) {
        int t = 0;
        while (lp[t] != l) ++t;
        return val[t];
    };
    int l = __ffsll((long long)mask) - 1;
    mask &= mask - 1ull;
    TreeFold f(l, leaf(l));
    while (mask) {
        l = __ffsll((long long)mask) - 1;
        mask &= mask - 1ull;
        f.push(l, leaf(l));
    }
    return f.finish();
}

// Stage 1 of a row (A entries a0 .. a0 + L, products base .. base + ip):
// product j = 64 r + lane (j < ip) reads B entry bi[r] of list lp[r], whose
// A value is sv[r] (VALS).  The k-th non-empty list marks its first product
// in a bitmap and fills table entry k; product j's list is then entry
// (non-empty lists starting at or before j) - 1.  Ends with the table read
// and its LDS free for reuse.
template <int R, bool VALS>
__device__ __forceinline__ void rm_products(const RowArgs& a, int a0, int L, int64_t base, int ip, RmLists& ls,
                                            int (&bi)[R], int (&lp)[R], double (&sv)[R]) {
    const int lane = lane_id();
    if (lane < R) ls.start[lane] = 0ull;
    int off = 0, bst = 0, len = 0;
    double sc = 0.0;
    if (lane < L) {
        off = (int)(a.ipA[a0 + lane] - base);
        len = (int)(a.ipA[a0 + lane + 1] - base) - off;
        bst = a.brpt[a.acol[a0 + lane]];
        if (VALS) sc = a.aval[a0 + lane];
    }
    const uint64_t nzl = __ballot(len > 0);
    wave_lds_sync();  // start words cleared
    if (len > 0) {
        atomicOr(&ls.start[off >> 6], 1ull << (off & 63));
        ls.ent[__popcll(nzl & ((1ull << lane) - 1ull))] = RmListEnt{off | (lane << 16), bst, sc};
    }
    wave_lds_sync();
    const uint64_t upto = ~0ull >> (63 - lane);  // bits 0..lane
    int before = 0;                               // non-empty lists starting before product 64 r
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = r * 64 + lane;
        lp[r] = 0;
        bi[r] = 0;
        sv[r] = 0.0;
        if (r * 64 >= ip) continue;  // uniform
        const uint64_t w = ls.start[r];
        if (j < ip) {
            const RmListEnt e = ls.ent[before + __popcll(w & upto) - 1];
            lp[r] = e.off_lp >> 16;
            bi[r] = e.bst + j - (e.off_lp & 0xffff);
            if (VALS) sv[r] = e.sc;
        }
        before += __popcll(w);
    }
    wave_lds_sync();  // the table is read before its area is reused
}

// Exclusive ranks of n words' popcounts into rank[0..n), consecutive lanes
// on consecutive words, at most NI steps of 64 words; returns the total.
template <int NI>
__device__ __forceinline__ int rank_words(const unsigned long long* bm, unsigned short* rank, int n) {
    const int lane = lane_id();
    int tot = 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        if (i * BRM_WAVE >= n) break;  // uniform
        const int w = i * BRM_WAVE + lane;
        const int pc = w < n ? __popcll(bm[w]) : 0;
        const int inc = wave_incl_scan_dpp(pc);
        if (w < n) rank[w] = (unsigned short)(tot + inc - pc);
        tot += __shfl(inc, BRM_WAVE - 1, BRM_WAVE);
    }
    return tot;
}

// Products of product indices 64 r + lane < ip: column cl[r] & CL_COL, list
// position cl[r] >> CL_LP, value prod[r]; B has at most RB_SPAN (flat) or
// RB2_SPAN (two-level) columns, nw words of them.
constexpr int CL_LP = 20;
constexpr int CL_COL = (1 << CL_LP) - 1;
static_assert(RB2_SPAN <= CL_COL + 1, "columns beside the list position");
template <int R, bool TWO>
__device__ void rmed_bitmap(const int (&cl)[R], const double (&prod)[R], int ip, int nw, int row, int64_t base,
                            const OutArgs& o, RbSmem<R, TWO>& sh) {
    const int lane = lane_id();
    int col[R];
#pragma unroll
    for (int r = 0; r < R; ++r) col[r] = cl[r] & CL_COL;
    RbBuckets<R, TWO>& B = sh.b;
    // only the words between the row's smallest and largest column are used
    int cmn = INT_MAX, cmx = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            cmn = min(cmn, col[r]);
            cmx = max(cmx, col[r]);
        }
    }
#pragma unroll
    for (int d = BRM_WAVE / 2; d > 0; d >>= 1) {
        cmn = min(cmn, __shfl_xor(cmn, d, BRM_WAVE));
        cmx = max(cmx, __shfl_xor(cmx, d, BRM_WAVE));
    }
    const int w0 = cmn >> 6, w1 = min((cmx >> 6) + 1, nw);
    const int cbase = TWO ? 0 : w0 * 64;  // flat: columns of the span are cbase + [0, 2^16)
    int nout = 0;
    int pos[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pos[r] = 0;
    if constexpr (!TWO) {
        constexpr int WPL = RB_WORDS / BRM_WAVE;  // bitmap words per lane and chunk, at most
        RbRanks& A = sh.a;
        // column - cbase of each product; lanes past the row hold a value no
        // chunk takes
        int bc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) bc[r] = r * 64 + lane < ip ? col[r] - cbase : (1 << 30);
        // the span in chunks of RB_WORDS words (two at most): bits, ranks and
        // the positions of the chunk's products; ranks carry across chunks
        auto chunk = [&](int wc) {
            const int nwc = min(RB_WORDS, w1 - w0 - wc);
#pragma unroll
            for (int i = 0; i < WPL; ++i) {
                const int w = i * BRM_WAVE + lane;
                if (w < nwc) A.bm[w] = 0ull;
            }
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const unsigned bit = (unsigned)(bc[r] - wc * 64);
                if (bit < (unsigned)(RB_WORDS * 64)) atomicOr(&A.bm[bit >> 6], 1ull << (bit & 63));
            }
            wave_lds_sync();
            const int base_rank = nout;
            nout += rank_words<WPL>(A.bm, A.wrank, nwc);
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const unsigned bit = (unsigned)(bc[r] - wc * 64);
                if (bit < (unsigned)(RB_WORDS * 64)) {
                    const int wd = (int)(bit >> 6);
                    pos[r] = base_rank + A.wrank[wd] + __popcll(A.bm[wd] & ((1ull << (bit & 63)) - 1ull));
                }
            }
            wave_lds_sync();  // read before the next chunk clears
        };
        if (w1 > w0) chunk(0);
        if (w1 - w0 > RB_WORDS) chunk(RB_WORDS);  // uniform
    } else {
        Rb2Ranks<R>& A = sh.a;
        constexpr int L1PL = RB2_L1 / BRM_WAVE;  // level-1 words per lane, at most
        const int n1 = (w1 - w0 + 63) >> 6;       // level-1 words in use
#pragma unroll
        for (int i = 0; i < L1PL; ++i) {
            const int w = i * BRM_WAVE + lane;
            if (w < n1) A.l1[w] = 0ull;
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r * 64 + lane < ip) {
                const int wid = (col[r] >> 6) - w0;
                atomicOr(&A.l1[wid >> 6], 1ull << (wid & 63));
            }
        }
        wave_lds_sync();
        const int n2 = rank_words<L1PL>(A.l1, A.l1rank, n1);  // level-2 words (words in use)
        wave_lds_sync();
        int kk[R];  // level-2 word of each product
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kk[r] = 0;
            if (r * 64 + lane < ip) {
                const int wid = (col[r] >> 6) - w0;
                kk[r] = A.l1rank[wid >> 6] + __popcll(A.l1[wid >> 6] & ((1ull << (wid & 63)) - 1ull));
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int q = i * BRM_WAVE + lane;
            if (q < n2) A.l2[q] = 0ull;
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r * 64 + lane < ip) atomicOr(&A.l2[kk[r]], 1ull << (col[r] & 63));
        wave_lds_sync();
        nout = rank_words<R>(A.l2, A.l2rank, n2);
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r * 64 + lane < ip)
                pos[r] = A.l2rank[kk[r]] + __popcll(A.l2[kk[r]] & ((1ull << (col[r] & 63)) - 1ull));
        }
        wave_lds_sync();
    }
    // ranks are dead from here on
    // bucket counters: 16 bits each, buckets 2i and 2i + 1 in word i
    const int ncw = (nout + 1) >> 1;
#pragma unroll
    for (int i = 0; i < R / 2; ++i)
        if (i * BRM_WAVE + lane < ncw) B.cnt[i * BRM_WAVE + lane] = 0u;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            atomicAdd(&B.cnt[pos[r] >> 1], 1u << ((pos[r] & 1) << 4));
            // every writer stores the same column
            if constexpr (TWO) B.oc[pos[r]] = col[r];
            else B.oc[pos[r]] = (unsigned short)(col[r] - cbase);
        }
    }
    wave_lds_sync();
    {
        int carry = 0;  // bucket offsets: a scan per step of 128 buckets, carried
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            if (i * BRM_WAVE >= ncw) break;  // uniform
            const int q2 = i * BRM_WAVE + lane;
            const unsigned w = q2 < ncw ? B.cnt[q2] : 0u;
            const int c0 = (int)(w & 0xffffu), t = c0 + (int)(w >> 16);
            const int inc = wave_incl_scan_dpp(t);
            const int ex = carry + inc - t;
            if (q2 < ncw) B.cnt[q2] = (unsigned)ex | ((unsigned)(ex + c0) << 16);
            carry += __shfl(inc, BRM_WAVE - 1, BRM_WAVE);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            const unsigned sft = (unsigned)(pos[r] & 1) << 4;
            const int slot = (int)((atomicAdd(&B.cnt[pos[r] >> 1], 1u << sft) >> sft) & 0xffffu);
            B.val[slot] = prod[r];
            B.lp[slot] = (unsigned char)(cl[r] >> CL_LP);
        }
    }
    wave_lds_sync();
    // counter q is now the end of bucket q
    auto bend = [&](int q) { return (int)((B.cnt[q >> 1] >> ((q & 1) << 4)) & 0xffffu); };
    int* dcol;
    double* dval;
    int64_t dcap;
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        dcol = o.ccol + cb;
        dval = o.cval + cb;
        dcap = o.crpt[row + 1] - cb;
    } else {
        dcol = o.ccol + base;
        dval = o.cval + base;
        dcap = ip;
    }
    for (int q = lane; q < nout; q += BRM_WAVE) {
        const int s = q ? bend(q - 1) : 0;
        const int m = bend(q) - s;
        double v;
        if (m == 1) {
            v = B.val[s];
        } else if (m == 2) {
            const double x = B.val[s], y = B.val[s + 1];
            v = B.lp[s] < B.lp[s + 1] ? fadd(x, y) : fadd(y, x);
        } else {
            v = rb_bucket_sum(B.lp + s, B.val + s, m);
        }
        if (q < dcap) {
            dcol[q] = cbase + (int)B.oc[q];
            dval[q] = v;
        }
    }
    if (lane == 0) {
        if (o.upper) o.rowlen[row] = nout;
        else if ((int64_t)nout != dcap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
    }
}

// Bitonic sort of the wavefront's first 64 rr keys (rr <= R, a power of
// two, uniform), element i = r * 64 + lane; registers r >= rr are untouched.
template <int R, typename KT>
__device__ __forceinline__ void wave_bitonic_sort(KT (&key)[R], int rr) {
    constexpr int N = 64 * R;
    const int lane = lane_id();
    const int nrt = 64 * rr;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
        if (k > nrt) break;
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int rp = r ^ (j >> 6);
                    if (r < rp && rp < rr) {
                        const int i = r * 64 + lane;
                        const bool asc = (i & k) == 0;
                        const KT x = key[r], y = key[rp];
                        const bool sw = asc ? (x > y) : (x < y);
                        key[r] = sw ? y : x;
                        key[rp] = sw ? x : y;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r >= rr) break;
                    const int i = r * 64 + lane;
                    const bool asc = (i & k) == 0;
                    const bool lower = (lane & j) == 0;
                    const KT x = key[r];
                    const KT y = shfl_xor_key(x, j);
                    const KT mn = x < y ? x : y;
                    const KT mx = x < y ? y : x;
                    key[r] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

// MODE 0: sort; 1: flat bitmap (B of at most RB_SPAN columns); 2: two-level
// bitmap (B of at most RB2_SPAN columns).
template <int R, typename KT, int MODE>
__global__ __launch_bounds__(BRM_WAVE) void k_rmed(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o) {
    using Key = RmKey<R, KT>;
    // one of the layouts, sized for the variant; the list table of stage 1
    // overlays the start of each
    __shared__ typename std::conditional<MODE == 0, RmSmem<R, KT>, RbSmem<R, MODE == 2>>::type shm;
    RmSmem<R, KT>& sm = *reinterpret_cast<RmSmem<R, KT>*>(&shm);
    RmLists& ls = *reinterpret_cast<RmLists*>(&shm);
    const int lane = lane_id();
    const int row = rows[blockIdx.x];
    const int a0 = a.arpt[row];
    const int L = a.arpt[row + 1] - a0;
    const int64_t base = a.ipA[a0];
    const int ip = (int)(a.ipA[a0 + L] - base);

    // ---- 1. list descriptors in lanes, products in registers
    int bi[R], lp[R];
    double sv[R];
    rm_products<R, true>(a, a0, L, base, ip, ls, bi, lp, sv);
    int col[R];
    double prod[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        col[r] = 0;
        prod[r] = 0.0;
        if (r * 64 + lane < ip) {
            col[r] = a.bcol[bi[r]];
            prod[r] = fmul(sv[r], a.bval[bi[r]]);
        }
    }

    // ---- bitmap ranks instead of a sort
    if constexpr (MODE != 0) {
        int cl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) cl[r] = col[r] | (lp[r] << CL_LP);
        rmed_bitmap<R, MODE == 2>(cl, prod, ip, (a.N + 63) >> 6, row, base, o, shm);
        return;
    }
    KT key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = r * 64 + lane;
        key[r] = j < ip ? Key::make(col[r], lp[r], j) : ~(KT)0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r * 64 < ip) sm.val[r * 64 + lane] = prod[r];

    // ---- 2. sort by (column, list position): only the registers the row uses
    int rr = 1;
    while (64 * rr < ip) rr <<= 1;
    wave_bitonic_sort<R, KT>(key, rr);

    // ---- 3. run heads; sums of the runs
#pragma unroll
    for (int r = 0; r < R; ++r) sm.key[r * 64 + lane] = key[r];
    __syncthreads();
    bool head[R];
    int nout = 0;
    int idx[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = r * 64 + lane;
        const uint32_t c = Key::col(key[r]);
        const bool valid = i < ip;
        const bool h = valid && (i == 0 || Key::col(sm.key[i - 1]) != c);
        head[r] = h;
        const uint64_t bal = __ballot(h);
        idx[r] = nout + __popcll(bal & ((1ull << lane) - 1ull));
        nout += __popcll(bal);
    }
    int* dcol;
    double* dval;
    int64_t dcap;
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        dcol = o.ccol + cb;
        dval = o.cval + cb;
        dcap = o.crpt[row + 1] - cb;
    } else {
        const int64_t cb = base;
        dcol = o.ccol + cb;
        dval = o.cval + cb;
        dcap = ip;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!head[r]) continue;
        const int i = r * 64 + lane;
        const uint32_t c = Key::col(key[r]);
        int prev = Key::lp(key[r]);
        double cur = sm.val[Key::j(key[r])];
        double lv[RM_LOGL];
        unsigned pend = 0u;
        for (int q = i + 1; q < ip; ++q) {
            const KT kq = sm.key[q];
            if (Key::col(kq) != c) break;
            const int l = Key::lp(kq);
            const int d = 31 - __clz(prev ^ l);
            double v = cur;
#pragma unroll
            for (int h = 0; h < RM_LOGL; ++h)
                if (h < d && ((pend >> h) & 1u)) v = fadd(lv[h], v);
#pragma unroll
            for (int h = 0; h < RM_LOGL; ++h)
                if (h == d) lv[h] = v;
            pend = (pend & ~((1u << d) - 1u)) | (1u << d);
            cur = sm.val[Key::j(kq)];
            prev = l;
        }
        double v = cur;
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if ((pend >> h) & 1u) v = fadd(lv[h], v);
        if (idx[r] < dcap) {
            dcol[idx[r]] = (int)c;
            dval[idx[r]] = v;
        }
    }
    if (lane == 0) {
        if (o.upper) o.rowlen[row] = nout;
        else if ((int64_t)nout != dcap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
    }
}

// Symbolic count of register-tier rows: the row's length is the number of
// distinct columns, the bits set in an LDS bitmap.
//   flat (B with at most RB_SPAN columns): a bitmap of CB_BITS columns, one
//     pass per CB_BITS of the row's column span (at most two);
//   two-level (B with at most RB2_SPAN columns): as the numeric variant's
//     ranks -- level 1 numbers the 64-column words in use, level 2 holds them.
constexpr int CB_WORDS = 512;
constexpr int CB_BITS = CB_WORDS * 64;
template <int R, bool TWO>
union RcSmem {
    RmLists l;
    unsigned long long bm[CB_WORDS];
    struct {
        unsigned long long l1[RB2_L1];
        unsigned short l1rank[RB2_L1];
        unsigned long long l2[TWO ? 64 * R : 1];
    } a;
};

template <int R, bool TWO>
__global__ __launch_bounds__(BRM_WAVE) void k_rmed_count(const int* __restrict__ rows, int nrows, RowArgs a,
                                                         int64_t* __restrict__ nnzrow) {
    __shared__ RcSmem<R, TWO> sh;
    const int lane = lane_id();
    const int row = rows[blockIdx.x];
    const int a0 = a.arpt[row];
    const int L = a.arpt[row + 1] - a0;
    const int64_t base = a.ipA[a0];
    const int ip = (int)(a.ipA[a0 + L] - base);
    int bi[R], lp[R];
    double sv[R];
    rm_products<R, false>(a, a0, L, base, ip, sh.l, bi, lp, sv);
    int col[R];
    int cmn = INT_MAX, cmx = -1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        col[r] = INT_MAX;
        if (r * 64 + lane < ip) {
            col[r] = a.bcol[bi[r]];
            cmn = min(cmn, col[r]);
            cmx = max(cmx, col[r]);
        }
    }
#pragma unroll
    for (int d = BRM_WAVE / 2; d > 0; d >>= 1) {
        cmn = min(cmn, __shfl_xor(cmn, d, BRM_WAVE));
        cmx = max(cmx, __shfl_xor(cmx, d, BRM_WAVE));
    }
    int cnt = 0;
    if constexpr (!TWO) {
        // columns < RB_SPAN: c0 stays far from overflow; no products, no pass
        for (int c0 = cmn & ~63; c0 <= cmx; c0 += CB_BITS) {  // uniform
            const int nwd = min(CB_WORDS, ((cmx - c0) >> 6) + 1);
#pragma unroll
            for (int i = 0; i < CB_WORDS / BRM_WAVE; ++i)
                if (i * BRM_WAVE + lane < nwd) sh.bm[i * BRM_WAVE + lane] = 0ull;
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const unsigned d = (unsigned)(col[r] - c0);
                if (d < (unsigned)CB_BITS) atomicOr(&sh.bm[d >> 6], 1ull << (d & 63));
            }
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < CB_WORDS / BRM_WAVE; ++i)
                if (i * BRM_WAVE + lane < nwd) cnt += __popcll(sh.bm[i * BRM_WAVE + lane]);
            wave_lds_sync();  // counted before the next pass clears
        }
    } else if (cmx >= 0) {  // uniform: the row has products
        constexpr int L1PL = RB2_L1 / BRM_WAVE;
        const int w0 = cmn >> 6;
        const int n1 = ((cmx >> 6) - w0 + 64) >> 6;  // level-1 words in use
#pragma unroll
        for (int i = 0; i < L1PL; ++i)
            if (i * BRM_WAVE + lane < n1) sh.a.l1[i * BRM_WAVE + lane] = 0ull;
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r * 64 + lane < ip) {
                const int wid = (col[r] >> 6) - w0;
                atomicOr(&sh.a.l1[wid >> 6], 1ull << (wid & 63));
            }
        }
        wave_lds_sync();
        const int n2 = rank_words<L1PL>(sh.a.l1, sh.a.l1rank, n1);
        wave_lds_sync();
        int kk[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kk[r] = 0;
            if (r * 64 + lane < ip) {
                const int wid = (col[r] >> 6) - w0;
                kk[r] = sh.a.l1rank[wid >> 6] + __popcll(sh.a.l1[wid >> 6] & ((1ull << (wid & 63)) - 1ull));
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (i * BRM_WAVE + lane < n2) sh.a.l2[i * BRM_WAVE + lane] = 0ull;
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r * 64 + lane < ip) atomicOr(&sh.a.l2[kk[r]], 1ull << (col[r] & 63));
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (i * BRM_WAVE + lane < n2) cnt += __popcll(sh.a.l2[i * BRM_WAVE + lane]);
    }
#pragma unroll
    for (int d = BRM_WAVE / 2; d > 0; d >>= 1) cnt += __shfl_xor(cnt, d, BRM_WAVE);
    if (lane == 0) nnzrow[row] = cnt;
}

template <int R>
hipError_t rmed_count_t(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s) {
    const int64_t ch = launch_chunk(BRM_WAVE);
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        if (a.N <= RB_SPAN) k_rmed_count<R, false><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
        else k_rmed_count<R, true><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
    }
    return hipGetLastError();
}

template <int R>
hipError_t rmed_t(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s) {
    const int64_t ch = launch_chunk(BRM_WAVE);
    const bool narrow = a.N <= 0 || (int64_t)a.N - 1 <= RmKey<R, uint32_t>::COL_MAX;
    // the bitmap variants pay for their larger LDS footprint on rows of more
    // than 128 products
    const bool bm_on = R >= 4 && params().wmed_bitmap == 1;
    const int mode = !bm_on ? 0 : a.N <= RB_SPAN ? 1 : a.N <= RB2_SPAN ? 2 : 0;
    // rows of more than 1024 products are binned only where a bitmap variant runs
    if (R > 16 && mode == 0) return hipErrorInvalidValue;
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        if constexpr (R >= 4) {
            if (mode == 1) {
                k_rmed<R, uint32_t, 1><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
                continue;
            }
            if (mode == 2) {
                k_rmed<R, uint32_t, 2><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
                continue;
            }
        }
        if constexpr (R <= 16) {
            if (narrow) k_rmed<R, uint32_t, 0><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
            else k_rmed<R, uint64_t, 0><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
        }
    }
    return hipGetLastError();
}

}  // namespace

int rmed_auto_cap(int N) {
    // The bitmap variants take rows of up to WMED_CAP_BITMAP products while
    // B has at most RB2_SPAN columns (rows of 1025-2048 products: 2,048
    // products per wavefront ran faster than the workgroup LDS tier).
    // Otherwise 1024-product rows sort 32-bit keys only while N fits beside
    // 16 bits of position and index; with 64-bit keys they run slower than
    // the workgroup LDS tier, so the register tier then stops at 512 products.
    if (params().wmed_bitmap == 1 && N <= RB2_SPAN) return WMED_CAP_BITMAP;
    return (int64_t)N - 1 <= RmKey<16, uint32_t>::COL_MAX ? 1024 : 512;
}

int rmed_max_cap(int N) {
    return params().wmed_bitmap == 1 && N <= RB2_SPAN ? WMED_CAP_BITMAP : WMED_CAP;
}

bool rmed_count_fits(int N) { return N <= RB2_SPAN; }

hipError_t launch_rmed_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    if (!rmed_count_fits(a.N)) return hipErrorInvalidValue;
    ProfScope ps_("k_rmed_count", s);
    if (cap <= 128) return rmed_count_t<2>(rows, n, a, nnzrow, s);
    if (cap <= 256) return rmed_count_t<4>(rows, n, a, nnzrow, s);
    if (cap <= 512) return rmed_count_t<8>(rows, n, a, nnzrow, s);
    if (cap <= 1024) return rmed_count_t<16>(rows, n, a, nnzrow, s);
    return rmed_count_t<32>(rows, n, a, nnzrow, s);
}

hipError_t launch_rmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    ProfScope ps_("k_rmed", s);
    if (cap <= 128) return rmed_t<2>(rows, n, a, o, s);
    if (cap <= 256) return rmed_t<4>(rows, n, a, o, s);
    if (cap <= 512) return rmed_t<8>(rows, n, a, o, s);
    if (cap <= 1024) return rmed_t<16>(rows, n, a, o, s);
    return rmed_t<32>(rows, n, a, o, s);
}

}  // namespace brm
