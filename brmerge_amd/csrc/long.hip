// Long-row tier: rows whose products do not fit one workgroup's LDS.
//
// The row's merge tree (aligned binary tree over A-entry positions, as in
// BRMerge-SpGEMM/inc/spgemm.h:639-707) is cut at its frontier: the maximal
// aligned subtrees whose products fit the LDS tile.  Each frontier subtree
// with at least two non-empty lists is merged in LDS by the leaf kernel
// (lds.hip) into a workspace buffer; a frontier subtree with one non-empty
// list is left in place as a "virtual" list (that B row, scaled on the fly).
// Every node above the frontier that holds two or more non-empty lists is a
// pair task at its height; heights are processed bottom-up, each height by
// four kernels over all its tasks:
//   len    -- input lengths and tile counts per task,
//   split  -- merge-path split of every 2048-element tile (global binary search),
//   count  -- per-tile output count (columns only, merged in LDS),
//   write  -- per-tile merge in LDS, staged and streamed out coalesced.
// Workspace storage alternates between two buffers by tree height (the
// reference's ping-pong), each node stored at its product offset in the row.
#include <limits.h>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int PL_NT = 256;

__device__ __forceinline__ int ceil_log2i(int x) { return x <= 1 ? 0 : 32 - __clz(x - 1); }
__device__ __forceinline__ int node_end(int L, int s, int h) {
    return (h >= 30) ? L : min(L, s + (1 << h));
}
__device__ __forceinline__ int64_t heap_idx(int H, int s, int h) {
    return ((int64_t)1 << (H - h)) - 1 + (s >> h);
}

struct NodeK {
    int kind;
    int q;
    int64_t ip;
};

__device__ NodeK node_kind(const int64_t* ipr, int L, int s, int h) {
    const int e = node_end(L, s, h);
    const int64_t ip = ipr[e] - ipr[s];
    if (ip == 0) return {SRC_EMPTY, 0, 0};
    const int q = upper_idx(ipr, s, e, ipr[s]);  // first non-empty list
    if (ipr[q + 1] - ipr[q] == ip) return {SRC_VIRTUAL, q, ip};
    return {SRC_STORED, -1, ip};
}

__device__ __forceinline__ bool fits(const int64_t* ipr, int L, int s, int h, int cap) {
    const int e = node_end(L, s, h);
    return (e - s) <= MED_LMAX && (ipr[e] - ipr[s]) <= cap;
}

__device__ Src make_src(const LongArgs& la, int a0, const int64_t* ipr, int L, int H, int s, int h,
                        int64_t wsb, int64_t slb) {
    const NodeK nk = node_kind(ipr, L, s, h);
    Src r;
    r.kind = nk.kind;
    r.buf = 0;
    r.off = 0;
    r.len = 0;
    r.slot = 0;
    r.scale = 0.0;
    if (nk.kind == SRC_VIRTUAL) {
        const int ai = a0 + nk.q;
        r.off = la.brpt[la.acol[ai]];
        r.len = nk.ip;
        r.scale = la.aval[ai];
    } else if (nk.kind == SRC_STORED) {
        r.buf = h & 1;
        r.off = wsb + (ipr[s] - ipr[0]);
        r.slot = slb + heap_idx(H, s, h);
    }
    return r;
}

__device__ void emit_task(const LongArgs& la, int h, const PairTask& T, int* err) {
    const unsigned long long idx = atomicAdd((unsigned long long*)&la.counters[CNT_TASK0 + h], 1ull);
    const int64_t cap = la.task_seg[h + 1] - la.task_seg[h];
    if ((int64_t)idx >= cap) {
        atomicOr(err, (int)DERR_CAPACITY);
        return;
    }
    la.tasks[la.task_seg[h] + (int64_t)idx] = T;
}

__global__ __launch_bounds__(PL_NT) void k_plan(LongArgs la, int cap, int* err) {
    const int r = blockIdx.x;
    const int row = la.rows[r];
    const int a0 = la.arpt[row];
    const int L = la.arpt[row + 1] - a0;
    const int64_t* ipr = la.ipA + a0;
    const int H = ceil_log2i(L);
    const int64_t wsb = la.ws_base[r], slb = la.slot_base[r];
    const int64_t final_off = la.final_off ? la.final_off[row] : ipr[0];
    const int hcap = min(H, MED_LOG_LMAX);

    for (int p = threadIdx.x; p < L; p += PL_NT) {
        int hf = 0;
        if (fits(ipr, L, p, 0, cap)) {
            while (hf < hcap && fits(ipr, L, p & ~((2 << hf) - 1), hf + 1, cap)) hf++;
        }
        const int s = p & ~((1 << hf) - 1);
        if (s != p) continue;
        const NodeK fk = node_kind(ipr, L, s, hf);
        if (fk.kind == SRC_STORED) {
            const unsigned long long li = atomicAdd((unsigned long long*)&la.counters[CNT_LEAF], 1ull);
            atomicAdd((unsigned long long*)&la.counters[CNT_LEAF_IP], (unsigned long long)fk.ip);
            LeafTask lt;
            lt.row = row;
            lt.ab = a0 + s;
            lt.ae = a0 + node_end(L, s, hf);
            lt.buf = hf & 1;
            lt.dst_off = wsb + (ipr[s] - ipr[0]);
            lt.slot = slb + heap_idx(H, s, hf);
            la.leaves[li] = lt;
        }
        if (hf == H) {  // the whole row is one frontier node: copy it to the destination
            PairTask T;
            T.L = make_src(la, a0, ipr, L, H, s, hf, wsb, slb);
            T.R = T.L;
            T.R.kind = SRC_EMPTY;
            T.row = row;
            T.final_dst = 1;
            T.dst_buf = 0;
            T.pad = 0;
            T.dst_off = final_off;
            T.slot = -1;
            emit_task(la, 1, T, err);
            continue;
        }
        for (int h = hf + 1; h <= H && (s & ((1 << h) - 1)) == 0; ++h) {
            const bool root = (h == H);
            const NodeK u = node_kind(ipr, L, s, h);
            if (u.kind != SRC_STORED && !root) continue;
            PairTask T;
            if (u.kind == SRC_STORED) {
                T.L = make_src(la, a0, ipr, L, H, s, h - 1, wsb, slb);
                const int s2 = s + (1 << (h - 1));
                if (s2 < L) {
                    T.R = make_src(la, a0, ipr, L, H, s2, h - 1, wsb, slb);
                } else {
                    T.R = T.L;
                    T.R.kind = SRC_EMPTY;
                }
            } else {  // the root holds a single non-empty list
                T.L = make_src(la, a0, ipr, L, H, s, h, wsb, slb);
                T.R = T.L;
                T.R.kind = SRC_EMPTY;
            }
            T.row = row;
            T.pad = 0;
            if (root) {
                T.final_dst = 1;
                T.dst_buf = 0;
                T.dst_off = final_off;
                T.slot = -1;
            } else {
                T.final_dst = 0;
                T.dst_buf = h & 1;
                T.dst_off = wsb + (ipr[s] - ipr[0]);
                T.slot = slb + heap_idx(H, s, h);
            }
            emit_task(la, h, T, err);
        }
    }
    if (threadIdx.x == 0)
        atomicMax((unsigned long long*)&la.counters[CNT_MAXH], (unsigned long long)(H < 1 ? 1 : H));
}

// ---------------------------------------------------------------- levels
struct WsView {
    int* col0;
    int* col1;
    double* val0;
    double* val1;
    int64_t* slots;
};

__device__ __forceinline__ int64_t src_len(const Src& s, const int64_t* slots) {
    return s.kind == SRC_EMPTY ? 0 : (s.kind == SRC_VIRTUAL ? s.len : slots[s.slot]);
}
__device__ __forceinline__ int src_col(const Src& s, const RowArgs& a, const WsView& w, int64_t x) {
    if (s.kind == SRC_VIRTUAL) return a.bcol[s.off + x];
    return (s.buf ? w.col1 : w.col0)[s.off + x];
}
__device__ __forceinline__ double src_val(const Src& s, const RowArgs& a, const WsView& w, int64_t x) {
    if (s.kind == SRC_VIRTUAL) return fmul(s.scale, a.bval[s.off + x]);
    return (s.buf ? w.val1 : w.val0)[s.off + x];
}

constexpr int LT_NT = LVL_NT;
constexpr int LT_CAP = LVL_TILE + 8;

__global__ void k_lvl_len(const PairTask* __restrict__ tasks, int64_t n, WsView w, int64_t* __restrict__ lens,
                          int64_t* __restrict__ ntile) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const PairTask T = tasks[t];
    const int64_t lL = src_len(T.L, w.slots), lR = src_len(T.R, w.slots);
    lens[2 * t] = lL;
    lens[2 * t + 1] = lR;
    const int64_t raw = lL + lR;
    ntile[t] = raw <= 0 ? 1 : (raw + LVL_TILE - 1) / LVL_TILE;
}

__global__ void k_lvl_split(const PairTask* __restrict__ tasks, int64_t n, const int64_t* __restrict__ toff,
                            int64_t ntiles, const int64_t* __restrict__ lens, RowArgs a, WsView w,
                            int64_t* __restrict__ split) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= ntiles) return;
    const int64_t t = upper_idx(toff, 0, (int)n, k);
    const int64_t d = (k - toff[t]) * (int64_t)LVL_TILE;
    const PairTask T = tasks[t];
    int64_t i = 0, j = 0;
    if (d > 0) {
        merge_split64([&](int64_t x) { return src_col(T.L, a, w, x); }, lens[2 * t],
                      [&](int64_t y) { return src_col(T.R, a, w, y); }, lens[2 * t + 1], d, &i, &j);
    }
    split[2 * k] = i;
    split[2 * k + 1] = j;
}

struct TileSmem {
    double vl[LT_CAP];
    double vr[LT_CAP];
    double ov[LT_CAP];
    int cl[LT_CAP];
    int cr[LT_CAP];
    int oc[LT_CAP];
    int si[LT_NT + 1];
    int sj[LT_NT + 1];
    int64_t red[LT_NT / BRM_WAVE + 1];
    int64_t misc[4];
};

// Loads this tile's two input ranges into LDS; returns (nl, nr) and task id.
template <bool VALS>
__device__ void load_tile(TileSmem& sm, const PairTask& T, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
                          const RowArgs& a, const WsView& w) {
    const int nl = (int)(i1 - i0), nr = (int)(j1 - j0);
    for (int x = threadIdx.x; x < nl; x += LT_NT) {
        sm.cl[x] = src_col(T.L, a, w, i0 + x);
        if (VALS) sm.vl[x] = src_val(T.L, a, w, i0 + x);
    }
    for (int y = threadIdx.x; y < nr; y += LT_NT) {
        sm.cr[y] = src_col(T.R, a, w, j0 + y);
        if (VALS) sm.vr[y] = src_val(T.R, a, w, j0 + y);
    }
}

__device__ void tile_splits(TileSmem& sm, int nl, int nr) {
    const int tot = nl + nr;
    const int c = (tot + LT_NT - 1) / LT_NT;
    for (int tt = threadIdx.x; tt <= LT_NT; tt += LT_NT) {
        const int d = min(tt * c, tot);
        int i, j;
        merge_split([&](int x) { return sm.cl[x]; }, nl, [&](int y) { return sm.cr[y]; }, nr, d, &i, &j);
        sm.si[tt] = i;
        sm.sj[tt] = j;
    }
}

__device__ __forceinline__ void tile_bounds(const int64_t* toff, int64_t n, const int64_t* lens,
                                            const int64_t* split, int64_t k, int64_t* t_out, int64_t* i0,
                                            int64_t* j0, int64_t* i1, int64_t* j1) {
    const int64_t t = upper_idx(toff, 0, (int)n, k);
    const int64_t last = toff[t + 1] - 1;
    *t_out = t;
    *i0 = split[2 * k];
    *j0 = split[2 * k + 1];
    if (k == last) {
        *i1 = lens[2 * t];
        *j1 = lens[2 * t + 1];
    } else {
        *i1 = split[2 * k + 2];
        *j1 = split[2 * k + 3];
    }
}

__global__ __launch_bounds__(LT_NT) void k_lvl_count(const PairTask* __restrict__ tasks, int64_t n,
                                                     const int64_t* __restrict__ toff, const int64_t* __restrict__ lens,
                                                     const int64_t* __restrict__ split, RowArgs a, WsView w,
                                                     int64_t* __restrict__ tcnt) {
    __shared__ TileSmem sm;
    const int64_t k = blockIdx.x;
    int64_t t, i0, j0, i1, j1;
    tile_bounds(toff, n, lens, split, k, &t, &i0, &j0, &i1, &j1);
    const PairTask T = tasks[t];
    load_tile<false>(sm, T, i0, i1, j0, j1, a, w);
    const int nl = (int)(i1 - i0), nr = (int)(j1 - j0);
    __syncthreads();
    tile_splits(sm, nl, nr);
    __syncthreads();
    int i = sm.si[threadIdx.x], j = sm.sj[threadIdx.x];
    const int ie = sm.si[threadIdx.x + 1], je = sm.sj[threadIdx.x + 1];
    int64_t cnt = 0;
    while (i < ie && j < je) {
        const int x = sm.cl[i], y = sm.cr[j];
        i += (x <= y);
        j += (y <= x);
        cnt++;
    }
    cnt += (ie - i) + (je - j);
    const int64_t tot = block_sum<LT_NT>(cnt, sm.red);
    if (threadIdx.x == 0) tcnt[k] = tot;
}

__global__ __launch_bounds__(LT_NT) void k_lvl_write(const PairTask* __restrict__ tasks, int64_t n,
                                                     const int64_t* __restrict__ toff, const int64_t* __restrict__ lens,
                                                     const int64_t* __restrict__ split, const int64_t* __restrict__ tout,
                                                     RowArgs a, WsView w, OutArgs o) {
    __shared__ TileSmem sm;
    const int64_t k = blockIdx.x;
    int64_t t, i0, j0, i1, j1;
    tile_bounds(toff, n, lens, split, k, &t, &i0, &j0, &i1, &j1);
    const PairTask T = tasks[t];
    load_tile<true>(sm, T, i0, i1, j0, j1, a, w);
    const int nl = (int)(i1 - i0), nr = (int)(j1 - j0);
    __syncthreads();
    tile_splits(sm, nl, nr);
    __syncthreads();
    int i = sm.si[threadIdx.x], j = sm.sj[threadIdx.x];
    const int ie = sm.si[threadIdx.x + 1], je = sm.sj[threadIdx.x + 1];
    // this thread's output position within the tile = merged count before (i, j)
    int64_t cnt = 0;
    {
        int ii = i, jj = j;
        while (ii < ie && jj < je) {
            const int x = sm.cl[ii], y = sm.cr[jj];
            ii += (x <= y);
            jj += (y <= x);
            cnt++;
        }
        cnt += (ie - ii) + (je - jj);
    }
    int64_t tile_tot;
    int o_ = (int)block_excl_scan<LT_NT>(cnt, sm.red, &tile_tot);
    while (i < ie && j < je) {
        const int x = sm.cl[i], y = sm.cr[j];
        if (x == y) {
            sm.oc[o_] = x;
            sm.ov[o_] = fadd(sm.vl[i], sm.vr[j]);
            i++;
            j++;
        } else if (x < y) {
            sm.oc[o_] = x;
            sm.ov[o_] = sm.vl[i];
            i++;
        } else {
            sm.oc[o_] = y;
            sm.ov[o_] = sm.vr[j];
            j++;
        }
        o_++;
    }
    for (; i < ie; ++i, ++o_) { sm.oc[o_] = sm.cl[i]; sm.ov[o_] = sm.vl[i]; }
    for (; j < je; ++j, ++o_) { sm.oc[o_] = sm.cr[j]; sm.ov[o_] = sm.vr[j]; }
    __syncthreads();

    const int64_t first = toff[t];
    const int64_t pre = tout[k] - tout[first];
    int* dc;
    double* dv;
    if (T.final_dst) {
        dc = o.ccol + T.dst_off;
        dv = o.cval + T.dst_off;
    } else {
        dc = (T.dst_buf ? w.col1 : w.col0) + T.dst_off;
        dv = (T.dst_buf ? w.val1 : w.val0) + T.dst_off;
    }
    int64_t wlim = tile_tot;
    if (T.final_dst && !o.upper) {  // Precise: never write past the row of C
        const int64_t room = o.crpt[T.row + 1] - o.crpt[T.row] - pre;
        wlim = room < 0 ? 0 : (room < wlim ? room : wlim);
    }
    for (int x = threadIdx.x; x < (int)wlim; x += LT_NT) {
        dc[pre + x] = sm.oc[x];
        dv[pre + x] = sm.ov[x];
    }
    if (threadIdx.x == 0 && k == toff[t + 1] - 1) {
        const int64_t len = tout[k + 1] - tout[first];
        if (T.final_dst) {
            if (!o.upper) {
                if (len != o.crpt[T.row + 1] - o.crpt[T.row]) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
            } else {
                o.rowlen[T.row] = len;
            }
        } else {
            w.slots[T.slot] = len;
        }
    }
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_plan(const LongArgs& la, int leaf_cap, int* err, hipStream_t s) {
    if (la.nrows <= 0) return hipSuccess;
    {
        ProfScope ps_("k_plan", s);
        k_plan<<<la.nrows, PL_NT, 0, s>>>(la, leaf_cap, err);
    }
    return hipGetLastError();
}

hipError_t launch_level_len(const PairTask* t, int64_t n, const LevelBufs& wb, int64_t* lens, int64_t* ntile,
                            hipStream_t s) {
    WsView w{wb.ws_col[0], wb.ws_col[1], wb.ws_val[0], wb.ws_val[1], wb.slots};
    {
        ProfScope ps_("k_lvl_len", s);
        k_lvl_len<<<nblk(n, 256), 256, 0, s>>>(t, n, w, lens, ntile);
    }
    return hipGetLastError();
}

hipError_t launch_level_split(const PairTask* t, int64_t n, const int64_t* toff, int64_t ntiles,
                              const int64_t* lens, const RowArgs& a, const LevelBufs& wb, int64_t* split,
                              hipStream_t s) {
    WsView w{wb.ws_col[0], wb.ws_col[1], wb.ws_val[0], wb.ws_val[1], wb.slots};
    {
        ProfScope ps_("k_lvl_split", s);
        k_lvl_split<<<nblk(ntiles, 256), 256, 0, s>>>(t, n, toff, ntiles, lens, a, w, split);
    }
    return hipGetLastError();
}

hipError_t launch_level_count(const PairTask* t, int64_t n, const int64_t* toff, int64_t ntiles,
                              const int64_t* lens, const int64_t* split, const RowArgs& a, const LevelBufs& wb,
                              int64_t* tcnt, hipStream_t s) {
    WsView w{wb.ws_col[0], wb.ws_col[1], wb.ws_val[0], wb.ws_val[1], wb.slots};
    {
        ProfScope ps_("k_lvl_count", s);
        k_lvl_count<<<(unsigned)ntiles, LT_NT, 0, s>>>(t, n, toff, lens, split, a, w, tcnt);
    }
    return hipGetLastError();
}

hipError_t launch_level_write(const PairTask* t, int64_t n, const int64_t* toff, int64_t ntiles,
                              const int64_t* lens, const int64_t* split, const int64_t* tout, const RowArgs& a,
                              const LevelBufs& wb, const OutArgs& o, hipStream_t s) {
    WsView w{wb.ws_col[0], wb.ws_col[1], wb.ws_val[0], wb.ws_val[1], wb.slots};
    {
        ProfScope ps_("k_lvl_write", s);
        k_lvl_write<<<(unsigned)ntiles, LT_NT, 0, s>>>(t, n, toff, lens, split, tout, a, w, o);
    }
    return hipGetLastError();
}

}  // namespace brm
