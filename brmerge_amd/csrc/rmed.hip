// Register tier for small medium rows: one wavefront per row, the row's
// products held in registers (R per lane, ip <= 64 R <= 1024, at most 64 lists).
//
//   1. every lane l < L holds list l's product offset, B row start and A
//      value, and writes the B index and list position of each of its
//      list's products to an LDS map; each product reads its entry, loads
//      its B entry and forms key = (column, list position, index) and
//      value = a * b (rounded, BRMerge-SpGEMM/inc/spgemm.h:621).  Keys are
//      32-bit when the column fits beside 6 bits of list position and
//      log2(64 R) bits of index (half the shuffles and compares), else 64-bit.
//   Sort variant:
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
//   Bitmap variant (B of at most RB_SPAN columns, R >= 8): the columns set
//   bits of an LDS bitmap over the row's column span, popcount ranks give
//   each product its output position, products are bucketed by position and
//   each bucket is folded in list-position order -- no sort.
// The symbolic count of these rows uses the wavefront hash kernel of lds.hip.
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

// Bitmap variant (B with at most RB_SPAN columns): the row's columns set
// bits of an LDS bitmap, popcount ranks give every product its
// output position, products are bucketed by position and each bucket is
// summed in list-position order -- O(ip) work instead of a sort.
constexpr int RB_WORDS = 1024;
constexpr int RB_SPAN = RB_WORDS * 64;
template <int R>
struct RbSmem {
    double val[64 * R];
    union {
        unsigned long long bm[RB_WORDS];
        struct {
            int cnt[64 * R];
            int outcol[64 * R];
        } c;
    } u;
    union {
        unsigned short wrank[RB_WORDS];
        unsigned short contrib[64 * R];  // list position << 10 | product index
    } w;
};
static_assert(64 * 16 <= 1024, "product index in 10 bits");

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

// Sum of one bucket (m >= 2 contributions from distinct lists) in the
// aligned tree over list positions: leaves are visited in increasing list
// position (the bits of `mask`), each found in the bucket by a scan; the
// fold keeps one pending left subtree per height (siblings added left +
// right, a node with one present child passes it through).
__device__ double rb_bucket_sum(const unsigned short* c, int m, const double* val) {
    unsigned long long mask = 0ull;
    for (int t = 0; t < m; ++t) mask |= 1ull << (c[t] >> 10);
    auto leaf = [&](int l) {
        int t = 0;
        while ((c[t] >> 10) != l) ++t;
        return val[c[t] & 1023];
    };
    int prev = __ffsll((long long)mask) - 1;
    mask &= mask - 1ull;
    double cur = leaf(prev);
    double lv[RM_LOGL];
    unsigned pend = 0u;
    while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1ull;
        const int d = 31 - __clz(prev ^ l);
        double v = cur;
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if (h < d && ((pend >> h) & 1u)) v = fadd(lv[h], v);
#pragma unroll
        for (int h = 0; h < RM_LOGL; ++h)
            if (h == d) lv[h] = v;
        pend = (pend & ~((1u << d) - 1u)) | (1u << d);
        cur = leaf(l);
        prev = l;
    }
    double v = cur;
#pragma unroll
    for (int h = 0; h < RM_LOGL; ++h)
        if ((pend >> h) & 1u) v = fadd(lv[h], v);
    return v;
}

// Columns index the bitmap directly (B has at most RB_SPAN columns, nw
// words of them in use).
template <int R, typename KT>
__device__ void rmed_bitmap(const KT (&key)[R], int ip, int nw, int row, int64_t base, const OutArgs& o,
                            RbSmem<R>& b) {
    using Key = RmKey<R, KT>;
    constexpr int WPL = RB_WORDS / BRM_WAVE;  // bitmap words per lane, at most
    const int lane = lane_id();
    // only the words between the row's smallest and largest column are used
    int cmn = INT_MAX, cmx = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            const int c = (int)Key::col(key[r]);
            cmn = min(cmn, c);
            cmx = max(cmx, c);
        }
    }
#pragma unroll
    for (int d = BRM_WAVE / 2; d > 0; d >>= 1) {
        cmn = min(cmn, __shfl_xor(cmn, d, BRM_WAVE));
        cmx = max(cmx, __shfl_xor(cmx, d, BRM_WAVE));
    }
    // words w0 + 64 i + lane (consecutive lanes on consecutive words)
    const int w0 = cmn >> 6, w1 = min((cmx >> 6) + 1, nw);
    const int nit = (w1 - w0 + BRM_WAVE - 1) / BRM_WAVE;
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        const int w = w0 + i * BRM_WAVE + lane;
        if (i < nit && w < w1) b.u.bm[w] = 0ull;
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            const int bit = (int)Key::col(key[r]);
            atomicOr(&b.u.bm[bit >> 6], 1ull << (bit & 63));
        }
    }
    wave_lds_sync();
    int nout = 0;  // ranks: a scan per step of 64 words, carried
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        if (i >= nit) break;  // uniform
        const int w = w0 + i * BRM_WAVE + lane;
        const int pc = w < w1 ? __popcll(b.u.bm[w]) : 0;
        const int inc = wave_incl_scan_dpp(pc);
        if (w < w1) b.w.wrank[w] = (unsigned short)(nout + inc - pc);
        nout += __shfl(inc, BRM_WAVE - 1, BRM_WAVE);
    }
    wave_lds_sync();
    int pos[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        pos[r] = 0;
        if (r * 64 + lane < ip) {
            const int bit = (int)Key::col(key[r]);
            const int wd = bit >> 6;
            pos[r] = b.w.wrank[wd] + __popcll(b.u.bm[wd] & ((1ull << (bit & 63)) - 1ull));
        }
    }
    wave_lds_sync();  // bitmap and ranks are dead from here on
#pragma unroll
    for (int i = 0; i < R; ++i)
        if (i * BRM_WAVE < nout) b.u.c.cnt[i * BRM_WAVE + lane] = 0;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            atomicAdd(&b.u.c.cnt[pos[r]], 1);
            b.u.c.outcol[pos[r]] = (int)Key::col(key[r]);  // every writer stores the same column
        }
    }
    wave_lds_sync();
    {
        int carry = 0;  // bucket offsets: a scan per step of 64 buckets, carried
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i * BRM_WAVE >= nout) break;  // uniform
            const int q = i * BRM_WAVE + lane;
            const int c = b.u.c.cnt[q];
            const int inc = wave_incl_scan_dpp(c);
            b.u.c.cnt[q] = carry + inc - c;
            carry += __shfl(inc, BRM_WAVE - 1, BRM_WAVE);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r * 64 + lane < ip) {
            const int slot = atomicAdd(&b.u.c.cnt[pos[r]], 1);
            b.w.contrib[slot] = (unsigned short)((Key::lp(key[r]) << 10) | (r * 64 + lane));
        }
    }
    wave_lds_sync();
    // cnt[q] is now the end of bucket q
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
        const int s = q ? b.u.c.cnt[q - 1] : 0;
        const int m = b.u.c.cnt[q] - s;
        const unsigned short* c = b.w.contrib + s;
        double v;
        if (m == 1) {
            v = b.val[c[0] & 1023];
        } else if (m == 2) {
            const bool first = (c[0] >> 10) < (c[1] >> 10);
            const double x = b.val[c[0] & 1023], y = b.val[c[1] & 1023];
            v = first ? fadd(x, y) : fadd(y, x);
        } else {
            v = rb_bucket_sum(c, m, b.val);
        }
        if (q < dcap) {
            dcol[q] = b.u.c.outcol[q];
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

template <int R, typename KT, bool BITMAP>
__global__ __launch_bounds__(BRM_WAVE) void k_rmed(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o) {
    using Key = RmKey<R, KT>;
    // one of the two layouts, sized for the variant (val first in both)
    __shared__ typename std::conditional<BITMAP, RbSmem<R>, RmSmem<R, KT>>::type shm;
    RmSmem<R, KT>& sm = *reinterpret_cast<RmSmem<R, KT>*>(&shm);
    const int lane = lane_id();
    const int row = rows[blockIdx.x];
    const int a0 = a.arpt[row];
    const int L = a.arpt[row + 1] - a0;
    const int64_t base = a.ipA[a0];
    const int ip = (int)(a.ipA[a0 + L] - base);

    // ---- 1. list descriptors in lanes, products in registers
    int off = INT_MAX, bst = 0, len = 0;
    double sc = 0.0;
    if (lane < L) {
        off = (int)(a.ipA[a0 + lane] - base);
        len = (int)(a.ipA[a0 + lane + 1] - base) - off;
        bst = a.brpt[a.acol[a0 + lane]];
        sc = a.aval[a0 + lane];
    }
    KT key[R];
    int bi[R], lp[R];
    double sv[R];
    {
        // every list lane writes its products' B index and list position to
        // LDS (areas unused until later: the bitmap area, or the keys and
        // values of the sort), one read each
        int* pbi;
        unsigned char* plp;
        if constexpr (BITMAP) {
            static_assert(64 * R * 5 <= (int)sizeof(RbSmem<R>::u), "product map fits the bitmap area");
            auto& b = *reinterpret_cast<RbSmem<R>*>(&shm);
            pbi = reinterpret_cast<int*>(&b.u);
            plp = reinterpret_cast<unsigned char*>(pbi + 64 * R);
        } else {
            pbi = reinterpret_cast<int*>(sm.key);
            plp = reinterpret_cast<unsigned char*>(sm.val);
        }
        for (int k = 0; k < len; ++k) {
            pbi[off + k] = bst + k;
            plp[off + k] = (unsigned char)lane;
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = r * 64 + lane;
            lp[r] = 0;
            bi[r] = 0;
            sv[r] = 0.0;
            if (r * 64 >= ip) continue;  // uniform
            if (j < ip) {
                bi[r] = pbi[j];
                lp[r] = plp[j];
            }
            sv[r] = __shfl(sc, lp[r], BRM_WAVE);
        }
        wave_lds_sync();  // the map is read before its areas are reused
    }
    double val[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = r * 64 + lane;
        key[r] = ~(KT)0;
        val[r] = 0.0;
        if (j < ip) {
            key[r] = Key::make(a.bcol[bi[r]], lp[r], j);
            val[r] = a.bval[bi[r]];
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r * 64 < ip) sm.val[r * 64 + lane] = fmul(sv[r], val[r]);

    // ---- B with at most RB_SPAN columns: bitmap ranks instead of a sort
    if (BITMAP) {
        wave_lds_sync();  // the values above are read by other lanes
        rmed_bitmap<R, KT>(key, ip, (a.N + 63) >> 6, row, base, o, *reinterpret_cast<RbSmem<R>*>(&shm));
        return;
    }

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

template <int R>
hipError_t rmed_t(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s) {
    const int64_t ch = launch_chunk(BRM_WAVE);
    const bool narrow = a.N <= 0 || (int64_t)a.N - 1 <= RmKey<R, uint32_t>::COL_MAX;
    // the bitmap variant pays for its larger LDS footprint on rows of more
    // than 256 products
    const bool bitmap = R >= 8 && a.N <= RB_SPAN;
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        if constexpr (R >= 8) {
            if (bitmap) {
                k_rmed<R, uint32_t, true><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
                continue;
            }
        }
        if (narrow) k_rmed<R, uint32_t, false><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
        else k_rmed<R, uint64_t, false><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
    }
    return hipGetLastError();
}

}  // namespace

int rmed_auto_cap(int N) {
    // 1024-product rows sort 32-bit keys only while N fits beside 16 bits
    // of position and index; with 64-bit keys they run slower than the
    // workgroup LDS tier, so the register tier then stops at 512 products.
    return (int64_t)N - 1 <= RmKey<16, uint32_t>::COL_MAX ? 1024 : 512;
}

hipError_t launch_rmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    ProfScope ps_("k_rmed", s);
    if (cap <= 128) return rmed_t<2>(rows, n, a, o, s);
    if (cap <= 256) return rmed_t<4>(rows, n, a, o, s);
    if (cap <= 512) return rmed_t<8>(rows, n, a, o, s);
    return rmed_t<16>(rows, n, a, o, s);
}

}  // namespace brm
