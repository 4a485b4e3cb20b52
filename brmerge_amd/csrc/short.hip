// Short-row tier: rows with at most 64 intermediate products, one
// wavefront per row (two per wavefront, 32 lanes each, for rows of at most
// 32 products), everything in registers.  Each lane holds one product
// (column, list position, value); short_merge.h ranks them by (column,
// position) and evaluates BRMerge's tree column-wise.  The symbolic variant
// counts distinct columns (first occurrence of each column in the rank order).
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"
#include "short_merge.h"

namespace brm {
namespace {

constexpr int SH_NT = 256;
constexpr int SH_WPB = SH_NT / BRM_WAVE;

// G = 64: one row per wavefront; G = 32: two rows per wavefront, each in its
// own half (every shuffle, ballot mask and rank stays inside the half).
template <bool NUMERIC, int G>
__global__ __launch_bounds__(SH_NT) void k_short(const int* __restrict__ rows, int nrows, int64_t wbase, RowArgs a,
                                                 OutArgs o, int64_t* __restrict__ nnzrow) {
    constexpr int RPW = BRM_WAVE / G;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint64_t gmask = G == BRM_WAVE ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const int wid = (blockIdx.x * SH_WPB + (int)(threadIdx.x / BRM_WAVE)) * RPW + lane / G;
    if (__ballot(wid < nrows) == 0ull) return;  // whole wavefront
    const bool has = wid < nrows;  // a group past the end takes part with an empty row
    int row = 0, a0 = 0, L = 0, ip = 0;
    int64_t base = 0;
    if (has) {
        row = rows[wid];
        a0 = a.arpt[row];
        L = a.arpt[row + 1] - a0;
        base = a.ipA[a0];
        ip = (int)(a.ipA[a0 + L] - base);
    }

    // list lanes (gl < L) load their list's product offset, B row start and
    // A value; product gl finds its list p (largest p with off[p] <= gl) by
    // a binary search over the group's lanes and takes p's descriptor by
    // shuffles.  Rows with more lists than lanes search ipA in memory.
    int off = INT_MAX, lbst = 0;
    double lav = 0.0;
    if (gl < L) {
        off = (int)(a.ipA[a0 + gl] - base);
        lbst = a.brpt[a.acol[a0 + gl]];
        if (NUMERIC) lav = a.aval[a0 + gl];
    }
    int p = 0;
#pragma unroll
    for (int step = G / 2; step >= 1; step >>= 1) {
        const int cand = p + step;
        const int oc = __shfl(off, gbase + (cand & (G - 1)), BRM_WAVE);
        if (cand < L && oc <= gl) p = cand;
    }
    const int src = gbase + (p & (G - 1));
    const int p_off = __shfl(off, src, BRM_WAVE);
    const int p_bst = __shfl(lbst, src, BRM_WAVE);
    const double p_av = __shfl(lav, src, BRM_WAVE);
    const bool act = gl < ip;
    int col = INT_MAX;
    double val = 0.0;
    if (act) {
        int bi;
        double av = p_av;
        if (L <= G) {
            bi = p_bst + (gl - p_off);
        } else {
            p = upper_idx(a.ipA + a0, 0, L, base + gl);
            const int ai = a0 + p;
            bi = a.brpt[a.acol[ai]] + (int)(base + gl - a.ipA[ai]);
            if (NUMERIC) av = a.aval[ai];
        }
        col = a.bcol[bi];
        if (NUMERIC) val = fmul(av, a.bval[bi]);
    }
    const ShortMerged m = short_merge<NUMERIC, G>(col, p, val, act, ip, gl, gbase, gmask, short_keys_fit<G>(a.N));
    if (!NUMERIC) {
        if (has && gl == 0) nnzrow[row] = (int64_t)m.nout;
        return;
    }
    const bool alive = m.alive;
    const int idx = m.idx, nout = m.nout;
    col = m.col;
    val = m.val;
    if (!has) return;
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        const int64_t cap = o.crpt[row + 1] - cb;
        if (alive && idx < cap) {
            o.ccol[cb + idx] = col;
            o.cval[cb + idx] = val;
        }
        if (gl == 0 && (int64_t)nout != cap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
    } else {
        const int64_t ob = o.slot ? (wbase + wid) * o.slot : base;
        if (alive) {
            o.ccol[ob + idx] = col;
            o.cval[ob + idx] = val;
        }
        if (gl == 0) o.rowlen[row] = nout;
    }
}

// C row of short row k of the list, from its slot at k * G
template <int G>
__global__ __launch_bounds__(SH_NT) void k_short_copy(const int* __restrict__ rows, int nrows, int64_t wbase,
                                                      const int* __restrict__ scol, const double* __restrict__ sval,
                                                      const int64_t* __restrict__ crpt, int* __restrict__ ccol,
                                                      double* __restrict__ cval) {
    constexpr int RPW = BRM_WAVE / G;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int wid = (blockIdx.x * SH_WPB + (int)(threadIdx.x / BRM_WAVE)) * RPW + lane / G;
    if (wid >= nrows) return;
    const int row = rows[wid];
    const int64_t cb = crpt[row];
    const int len = (int)(crpt[row + 1] - cb);
    const int64_t sb = (wbase + wid) * G;
    if (gl < len) {
        ccol[cb + gl] = scol[sb + gl];
        cval[cb + gl] = sval[sb + gl];
    }
}

template <bool NUMERIC, int G>
void short_launch(const int* rows, int n, const RowArgs& a, const OutArgs& o, int64_t* nnzrow, hipStream_t s) {
    constexpr int RPB = SH_WPB * (BRM_WAVE / G);  // rows per block
    const int64_t ch = launch_chunk(G);
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        k_short<NUMERIC, G><<<(c + RPB - 1) / RPB, SH_NT, 0, s>>>(rows + b, c, b, a, o, nnzrow);
    }
}

}  // namespace

hipError_t launch_short_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s,
                              int group) {
    if (n <= 0) return hipSuccess;
    OutArgs o{};
    {
        ProfScope ps_("k_short_count", s);
        if (group == 32) short_launch<false, 32>(rows, n, a, o, nnzrow, s);
        else short_launch<false, 64>(rows, n, a, o, nnzrow, s);
    }
    return hipGetLastError();
}

hipError_t launch_short_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                                int group) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_short_numeric", s);
        if (group == 32) short_launch<true, 32>(rows, n, a, o, nullptr, s);
        else short_launch<true, 64>(rows, n, a, o, nullptr, s);
    }
    return hipGetLastError();
}

namespace {
template <int G>
void short_copy_t(const int* rows, int n, const int* scol, const double* sval, const int64_t* crpt, int* ccol,
                  double* cval, hipStream_t s) {
    constexpr int RPB = SH_WPB * (BRM_WAVE / G);
    const int64_t ch = launch_chunk(G);
    for (int64_t b = 0; b < n; b += ch) {
        const int c = (int)std::min<int64_t>(ch, n - b);
        k_short_copy<G><<<(c + RPB - 1) / RPB, SH_NT, 0, s>>>(rows + b, c, b, scol, sval, crpt, ccol, cval);
    }
}
}  // namespace

hipError_t launch_short_copy(const int* rows, int n, const int* scol, const double* sval, const int64_t* crpt,
                             int* ccol, double* cval, hipStream_t s, int group) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_short_copy", s);
        if (group == 32) short_copy_t<32>(rows, n, scol, sval, crpt, ccol, cval, s);
        else short_copy_t<64>(rows, n, scol, sval, crpt, ccol, cval, s);
    }
    return hipGetLastError();
}

}  // namespace brm
