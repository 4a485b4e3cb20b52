// Register tier for small medium rows: one wavefront per row, the row's
// products held in registers (R per lane, ip <= 64 R <= 1024, at most 64 lists).
//
//   1. every lane l < L holds list l's product offset, B row start and A
//      value; each product locates its list with a shuffle binary search,
//      loads its B entry and forms key = (column, list position, index)
//      and value = a * b (rounded, BRMerge-SpGEMM/inc/spgemm.h:621);
//   2. a bitonic sort over the 64 R keys (cross-lane steps with shuffles,
//      intra-lane steps between registers) orders the products by column,
//      and within a column by list position;
//   3. the sorted keys and the values go to LDS; the first product of every
//      column run sums its run as the aligned binary tree over list
//      positions does (siblings added left + right, a node with one present
//      child passes it through, spgemm.h:639-707) -- a single pass that
//      keeps one pending left subtree per height;
//   4. run heads are compacted with ballots and written to C (Precise), or
//      to C-bar with the row length (Upper).  (The symbolic count of these
//      rows uses the wavefront hash kernel of lds.hip.)
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int RM_LOGL = 6;  // list positions < 64: six tree levels

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
    const int lo = __shfl_xor((int)(uint32_t)x, m, BRM_WAVE);
    const int hi = __shfl_xor((int)(uint32_t)(x >> 32), m, BRM_WAVE);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <int R>
struct RmSmem {
    uint64_t key[64 * R];
    double val[64 * R];
};

// Bitonic sort of the wavefront's first 64 rr keys (rr <= R, a power of
// two, uniform), element i = r * 64 + lane; registers r >= rr are untouched.
template <int R>
__device__ __forceinline__ void wave_bitonic_sort(uint64_t (&key)[R], int rr) {
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
                        const uint64_t x = key[r], y = key[rp];
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
                    const uint64_t x = key[r];
                    const uint64_t y = shfl_xor_u64(x, j);
                    const uint64_t mn = x < y ? x : y;
                    const uint64_t mx = x < y ? y : x;
                    key[r] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

template <int R>
__global__ __launch_bounds__(BRM_WAVE) void k_rmed(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o) {
    __shared__ RmSmem<R> sm;
    const int lane = lane_id();
    const int row = rows[blockIdx.x];
    const int a0 = a.arpt[row];
    const int L = a.arpt[row + 1] - a0;
    const int64_t base = a.ipA[a0];
    const int ip = (int)(a.ipA[a0 + L] - base);

    // ---- 1. list descriptors in lanes, products in registers
    int off = INT_MAX, bst = 0;
    double sc = 0.0;
    if (lane < L) {
        off = (int)(a.ipA[a0 + lane] - base);
        bst = a.brpt[a.acol[a0 + lane]];
        sc = a.aval[a0 + lane];
    }
    uint64_t key[R];
    int bi[R], lp[R];
    double sv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = r * 64 + lane;
        int p = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {  // largest list p with off[p] <= j
            const int cand = p + step;
            const int oc = __shfl(off, cand & (BRM_WAVE - 1), BRM_WAVE);
            if (cand < L && oc <= j) p = cand;
        }
        const int o_p = __shfl(off, p, BRM_WAVE);
        const int b_p = __shfl(bst, p, BRM_WAVE);
        sv[r] = __shfl(sc, p, BRM_WAVE);
        lp[r] = p;
        bi[r] = b_p + (j - o_p);
    }
    double val[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = r * 64 + lane;
        key[r] = ~0ull;
        val[r] = 0.0;
        if (j < ip) {
            const int c = a.bcol[bi[r]];
            key[r] = ((uint64_t)(uint32_t)c << 32) | ((uint64_t)lp[r] << 16) | (uint64_t)j;
            val[r] = a.bval[bi[r]];
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) sm.val[r * 64 + lane] = fmul(sv[r], val[r]);

    // ---- 2. sort by (column, list position): only the registers the row uses
    int rr = 1;
    while (64 * rr < ip) rr <<= 1;
    wave_bitonic_sort<R>(key, rr);

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
        const uint32_t c = (uint32_t)(key[r] >> 32);
        const bool valid = i < ip;
        const bool h = valid && (i == 0 || (uint32_t)(sm.key[i - 1] >> 32) != c);
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
        const uint32_t c = (uint32_t)(key[r] >> 32);
        int prev = (int)((key[r] >> 16) & 0xFFFFu);
        double cur = sm.val[(int)(key[r] & 0xFFFFu)];
        double lv[RM_LOGL];
        unsigned pend = 0u;
        for (int q = i + 1; q < ip; ++q) {
            const uint64_t kq = sm.key[q];
            if ((uint32_t)(kq >> 32) != c) break;
            const int l = (int)((kq >> 16) & 0xFFFFu);
            const int d = 31 - __clz(prev ^ l);
            double v = cur;
#pragma unroll
            for (int h = 0; h < RM_LOGL; ++h)
                if (h < d && ((pend >> h) & 1u)) v = fadd(lv[h], v);
#pragma unroll
            for (int h = 0; h < RM_LOGL; ++h)
                if (h == d) lv[h] = v;
            pend = (pend & ~((1u << d) - 1u)) | (1u << d);
            cur = sm.val[(int)(kq & 0xFFFFu)];
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
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        k_rmed<R><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, o);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_rmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    ProfScope ps_("k_rmed", s);
    if (cap <= 128) return rmed_t<2>(rows, n, a, o, s);
    if (cap <= 256) return rmed_t<4>(rows, n, a, o, s);
    if (cap <= 512) return rmed_t<8>(rows, n, a, o, s);
    return rmed_t<16>(rows, n, a, o, s);
}

}  // namespace brm
