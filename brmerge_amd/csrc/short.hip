// Short-row tier: rows with at most 64 intermediate products, one
// wavefront per row, everything in registers.
//
// Each lane holds one product (column, list position, value).  A rank sort
// by (column, position) over the wavefront (v_readlane broadcasts) places the
// products in column order with equal columns adjacent and ordered by
// position; ds_permute moves every lane's product to its rank.  The merge tree
// of BRMerge (BRMerge-SpGEMM/inc/spgemm.h:639-707: lists paired (2j, 2j+1)
// level by level, odd list carried) is then evaluated column-wise: at level l
// two live entries of the same column combine iff their positions share the
// parent node (pos >> (l+1)); the left one keeps the sum, the right one dies.
// That is exactly the reference's value rule -- an aligned binary tree over
// list positions in which a node with one present child passes it through --
// so the sums are bit-identical.  The symbolic variant counts distinct
// columns (first occurrence of each column in the rank order).
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int SH_NT = 256;
constexpr int SH_WPB = SH_NT / BRM_WAVE;

__device__ __forceinline__ int permute_i32(int dst_lane, int x) {
    return __builtin_amdgcn_ds_permute(dst_lane << 2, x);
}
__device__ __forceinline__ double permute_f64(int dst_lane, double x) {
    int lo = __builtin_amdgcn_ds_permute(dst_lane << 2, __double2loint(x));
    int hi = __builtin_amdgcn_ds_permute(dst_lane << 2, __double2hiint(x));
    return __hiloint2double(hi, lo);
}

// Locates product `lane` of a row: returns its list position (A entry - a0).
// Uniform control flow across the wavefront (shuffles inside).
__device__ __forceinline__ int locate_product(const RowArgs& a, int a0, int L, int64_t base, int ip, int lane) {
    if (L <= BRM_WAVE) {
        int off = lane < L ? (int)(a.ipA[a0 + lane] - base) : INT_MAX;
        int lo = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            int cand = lo + step;
            int oc = __shfl(off, cand & (BRM_WAVE - 1), BRM_WAVE);
            if (cand < L && oc <= lane) lo = cand;
        }
        return lo;
    }
    if (lane >= ip) return 0;
    return upper_idx(a.ipA + a0, 0, L, base + lane);
}

template <bool NUMERIC>
__global__ __launch_bounds__(SH_NT) void k_short(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o,
                                                 int64_t* __restrict__ nnzrow) {
    const int wid = blockIdx.x * SH_WPB + threadIdx.x / BRM_WAVE;
    if (wid >= nrows) return;  // whole wavefront
    const int lane = lane_id();
    const int row = rows[wid];
    const int a0 = a.arpt[row];
    const int L = a.arpt[row + 1] - a0;
    const int64_t base = a.ipA[a0];
    const int ip = (int)(a.ipA[a0 + L] - base);

    const int p = locate_product(a, a0, L, base, ip, lane);
    const bool act = lane < ip;
    int col = INT_MAX;
    double val = 0.0;
    if (act) {
        const int ai = a0 + p;
        const int bi = a.brpt[a.acol[ai]] + (int)(base + lane - a.ipA[ai]);
        col = a.bcol[bi];
        if (NUMERIC) val = fmul(a.aval[ai], a.bval[bi]);
    }
    const uint64_t key = act ? (((uint64_t)(uint32_t)col << 32) | (uint32_t)p) : ~0ull;
    const uint32_t klo_me = (uint32_t)key, khi_me = (uint32_t)(key >> 32);

    int rank = 0;
    bool dup = false;
    for (int k = 0; k < ip; ++k) {
        uint32_t klo = (uint32_t)__builtin_amdgcn_readlane((int)klo_me, k);
        uint32_t khi = (uint32_t)__builtin_amdgcn_readlane((int)khi_me, k);
        uint64_t kk = ((uint64_t)khi << 32) | klo;
        bool lt = kk < key;
        rank += lt;
        dup |= lt && (khi == khi_me);
    }

    if (!NUMERIC) {
        uint64_t first = __ballot(act && !dup);
        if (lane == 0) nnzrow[row] = (int64_t)__popcll(first);
        return;
    }

    const int dst = act ? rank : lane;
    col = permute_i32(dst, col);
    int pos = permute_i32(dst, p);
    val = permute_f64(dst, val);
    bool alive = lane < ip;

    const int nlev = L <= 1 ? 0 : 32 - __clz(L - 1);
    for (int l = 0; l < nlev; ++l) {
        const uint64_t am = __ballot(alive);
        const uint64_t above = (lane == BRM_WAVE - 1) ? 0ull : (am & (~0ull << (lane + 1)));
        const int nxt = above ? (__ffsll((long long)above) - 1) : lane;
        const int ncol = __shfl(col, nxt, BRM_WAVE);
        const int npos = __shfl(pos, nxt, BRM_WAVE);
        const double nval = __shfl(val, nxt, BRM_WAVE);
        const bool same = alive && above != 0ull && ncol == col;
        if (__ballot(same) == 0ull) break;  // no column left to combine
        const bool comb = same && ((npos >> (l + 1)) == (pos >> (l + 1)));
        const uint64_t cm = __ballot(comb);
        const uint64_t below = am & ((1ull << lane) - 1ull);
        const int prv = below ? (63 - __clzll((long long)below)) : -1;
        const bool dies = prv >= 0 && ((cm >> prv) & 1ull);
        if (comb) val = fadd(val, nval);
        alive = alive && !dies;
    }

    const uint64_t am = __ballot(alive);
    const int idx = __popcll(am & ((1ull << lane) - 1ull));
    const int nout = __popcll(am);
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        const int64_t cap = o.crpt[row + 1] - cb;
        if (alive && idx < cap) {
            o.ccol[cb + idx] = col;
            o.cval[cb + idx] = val;
        }
        if (lane == 0 && (int64_t)nout != cap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
    } else {
        if (alive) {
            o.ccol[base + idx] = col;
            o.cval[base + idx] = val;
        }
        if (lane == 0) o.rowlen[row] = nout;
    }
}

}  // namespace

hipError_t launch_short_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    OutArgs o{};
    {
        ProfScope ps_("k_short_count", s);
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            k_short<false><<<(c + SH_WPB - 1) / SH_WPB, SH_NT, 0, s>>>(rows + b, c, a, o, nnzrow);
        }
    }
    return hipGetLastError();
}

hipError_t launch_short_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_short_numeric", s);
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            k_short<true><<<(c + SH_WPB - 1) / SH_WPB, SH_NT, 0, s>>>(rows + b, c, a, o, nullptr);
        }
    }
    return hipGetLastError();
}

}  // namespace brm
