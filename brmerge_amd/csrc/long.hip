// Long-row tier: rows whose products do not fit one workgroup's LDS.
//
// The row's merge tree (the aligned binary tree over A-entry positions of
// BRMerge-SpGEMM/inc/spgemm.h:639-707) is cut at its frontier: the maximal
// aligned subtrees whose products fit an LDS tile.  A frontier subtree with
// two or more non-empty lists is merged in LDS by the leaf kernel (lds.hip)
// into a workspace buffer; one with a single non-empty list stays in place
// as a "virtual" list (that B row, scaled on the fly).
//
// Above the frontier the tree is merged pass_bits levels per pass: pass j
// materialises every node at height h_j = min(j * pass_bits, H) from its
// children -- the highest materialised nodes below it (frontier nodes or
// pass j-1 nodes), at most 2^pass_bits of them.  (pass_bits is 5 for this
// file's merge-path engine, 6 for the bitmap engine of bmerge.hip, which
// also runs without a frontier: every list is a child of its pass-1 node.)  One workgroup streams a
// node's children through LDS in rounds: every child contributes a window
// proportional to what it has left, and all elements up to the smallest
// window-end column are merged (tree order, in LDS) and written out.  Every
// contribution to a column <= that bound is present in the round, so each
// column is summed by the reference's bracketing, bit for bit.
//
// Workspace buffers alternate by pass: pass j writes buffer j&1 at the node's
// product offset in the row and reads buffer (j-1)&1.
#include <limits.h>

#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "brm_device.h"
#include "kernels.h"
#include "lds_merge.h"

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
__device__ __forceinline__ int pass_height(int j, int H, int pb) { return min(j * pb, H); }

// A row's element offsets into workspace buffers 0 and 1 (each buffer holds
// only the rows that store nodes in it).
struct WsBase {
    int64_t b0, b1;
    __device__ int64_t at(int buf) const { return buf ? b1 : b0; }
};

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

struct Fit {
    int cap;       // products per LDS subtree
    int lmax;      // lists per LDS subtree
    int log_lmax;
};

__device__ __forceinline__ bool fits(const int64_t* ipr, int L, int s, int h, const Fit& f) {
    const int e = node_end(L, s, h);
    return (e - s) <= f.lmax && (ipr[e] - ipr[s]) <= f.cap;
}

// Height of the frontier node that contains position p.
__device__ int frontier_height(const int64_t* ipr, int L, int H, int p, const Fit& f) {
    const int hcap = min(H, f.log_lmax);
    int hf = 0;
    if (fits(ipr, L, p, 0, f)) {
        while (hf < hcap && fits(ipr, L, p & ~((2 << hf) - 1), hf + 1, f)) hf++;
    }
    return hf;
}

__device__ Src make_src(const LongArgs& la, int a0, const int64_t* ipr, int L, int H, int s, int h, int buf,
                        WsBase wsb, int64_t slb) {
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
        r.buf = buf;
        r.off = wsb.at(buf) + (ipr[s] - ipr[0]);
        r.slot = slb + heap_idx(H, s, h);
    }
    return r;
}

// Emits the group task of node (s, hj), the pass-j node (or the root).
__device__ void emit_group(const LongArgs& la, int row, int a0, const int64_t* ipr, int L, int H, int j, int s,
                           int hj, int hprev, const Fit& cap, WsBase wsb, int64_t slb, int64_t final_off, int* err) {
    const bool root = (hj == H);
    const NodeK u = node_kind(ipr, L, s, hj);
    if (u.kind == SRC_EMPTY || (u.kind == SRC_VIRTUAL && !root)) return;
    const int e = node_end(L, s, hj);
    // count, then reserve, then write the children
    int nc = 0;
    for (int q = s; q < e;) {
        const int hq = frontier_height(ipr, L, H, q, cap);
        const int ch = hq >= hprev ? hq : hprev;
        if (node_kind(ipr, L, q, ch).kind != SRC_EMPTY) nc++;
        q += 1 << ch;
    }
    const unsigned long long cb = atomicAdd((unsigned long long*)&la.counters[CNT_CHILD], (unsigned long long)nc);
    if ((int64_t)cb + nc > la.child_cap || nc > (1 << la.pass_bits)) {
        atomicOr(err, (int)DERR_CAPACITY);
        return;
    }
    int k = 0;
    for (int q = s; q < e;) {
        const int hq = frontier_height(ipr, L, H, q, cap);
        const int ch = hq >= hprev ? hq : hprev;
        const Src src = make_src(la, a0, ipr, L, H, q, ch, (j - 1) & 1, wsb, slb);
        if (src.kind != SRC_EMPTY) {
            Child c;
            c.src = src;
            // every child starts on a multiple of 2^hprev: dropping those
            // low bits keeps the pairing and skips levels with no pairs
            c.lid = (q - s) >> hprev;
            c.pad = 0;
            la.children[cb + k] = c;
            k++;
        }
        q += 1 << ch;
    }
    GroupTask T;
    T.row = row;
    T.nchild = nc;
    T.cbase = (int64_t)cb;
    if (root) {
        T.final_dst = 1;
        T.dst_buf = 0;
        T.dst_off = final_off;
        T.slot = -1;
    } else {
        T.final_dst = 0;
        T.dst_buf = j & 1;
        T.dst_off = wsb.at(j & 1) + (ipr[s] - ipr[0]);
        T.slot = slb + heap_idx(H, s, hj);
    }
    const unsigned long long idx = atomicAdd((unsigned long long*)&la.counters[CNT_TASK0 + j], 1ull);
    if ((int64_t)idx >= la.task_seg[j + 1] - la.task_seg[j]) {
        atomicOr(err, (int)DERR_CAPACITY);
        return;
    }
    la.tasks[la.task_seg[j] + (int64_t)idx] = T;
}

// Without a frontier (frontier_cap 0: every non-empty list is a child in
// place) the children of pass-j node (s, h_j) are exactly the non-empty
// nodes at height h_{j-1} inside it, at most PL_NT of them: the workgroup
// emits each node's children together, thread t taking child slot t, with
// a block scan for their order.
// Nodes of fan-in F = 2^(hj - hp) <= 2^pass_bits: the workgroup plans
// PL_NT / F nodes per step, one child slot per thread, with one scan and one
// reservation of child entries for all of them.
__device__ void plan_no_frontier(const LongArgs& la, int row, int a0, const int64_t* ipr, int L, int H, int J,
                                 WsBase wsb, int64_t slb, int64_t final_off, int* err) {
    __shared__ int64_t sm[PL_NT / BRM_WAVE + 1];
    __shared__ int64_t cbase;
    __shared__ int64_t lead[PL_NT + 1];  // child prefix at each node's first slot; [G]: the step's total
    const int t = threadIdx.x;
    const int pb = la.pass_bits;
    for (int j = 1; j <= J; ++j) {
        const int hj = pass_height(j, H, pb);
        const int hp = j == 1 ? 0 : pass_height(j - 1, H, pb);
        const bool root = hj == H;
        const int nnodes = root ? 1 : ((L - 1) >> hj) + 1;
        const int fb = hj - hp;      // log2 of the child slots per node
        const int G = PL_NT >> fb;   // nodes per step
        const int k = t >> fb;       // this thread's node in the step
        const int c = t & ((1 << fb) - 1);
        for (int nd0 = 0; nd0 < nnodes; nd0 += G) {
            const int nd = nd0 + k;
            bool live = false;
            int s = 0, e = 0;
            if (nd < nnodes) {
                s = nd << hj;
                const NodeK u = node_kind(ipr, L, s, hj);
                live = !(u.kind == SRC_EMPTY || (u.kind == SRC_VIRTUAL && !root));
                e = node_end(L, s, hj);
            }
            const int q = s + (c << hp);
            Src src{};
            src.kind = SRC_EMPTY;
            if (live && q < e) src = make_src(la, a0, ipr, L, H, q, hp, (j - 1) & 1, wsb, slb);
            const int64_t has = src.kind != SRC_EMPTY ? 1 : 0;
            int64_t nc;
            const int64_t rank = block_excl_scan<PL_NT>(has, sm, &nc);
            if (c == 0) lead[k] = rank;
            if (t == 0) {
                lead[G] = nc;
                const unsigned long long cb =
                    atomicAdd((unsigned long long*)&la.counters[CNT_CHILD], (unsigned long long)nc);
                cbase = (int64_t)cb;
                if ((int64_t)cb + nc > la.child_cap) atomicOr(err, (int)DERR_CAPACITY);
            }
            __syncthreads();
            const int64_t cb = cbase;
            if (cb + nc <= la.child_cap) {  // uniform
                if (has) {
                    Child ch;
                    ch.src = src;
                    ch.lid = c;  // (q - s) >> hp
                    ch.pad = 0;
                    la.children[cb + rank] = ch;
                }
                if (c == 0 && live) {
                    GroupTask T;
                    T.row = row;
                    T.nchild = (int)(lead[k + 1] - lead[k]);
                    T.cbase = cb + lead[k];
                    if (root) {
                        T.final_dst = 1;
                        T.dst_buf = 0;
                        T.dst_off = final_off;
                        T.slot = -1;
                    } else {
                        T.final_dst = 0;
                        T.dst_buf = j & 1;
                        T.dst_off = wsb.at(j & 1) + (ipr[s] - ipr[0]);
                        T.slot = slb + heap_idx(H, s, hj);
                    }
                    const unsigned long long idx =
                        atomicAdd((unsigned long long*)&la.counters[CNT_TASK0 + j], 1ull);
                    if ((int64_t)idx >= la.task_seg[j + 1] - la.task_seg[j]) {
                        atomicOr(err, (int)DERR_CAPACITY);
                    } else {
                        la.tasks[la.task_seg[j] + (int64_t)idx] = T;
                        if (la.task_key)
                            la.task_key[la.task_seg[j] + (int64_t)idx] =
                                (uint32_t)min(ipr[e] - ipr[s], (int64_t)UINT32_MAX);
                    }
                }
            }
            __syncthreads();  // cbase and lead are rewritten by the next step
        }
    }
}

__global__ __launch_bounds__(PL_NT) void k_plan(LongArgs la, int r0, int* err) {
    const Fit cap{la.frontier_cap, la.lmax, la.log_lmax};
    const int r = r0 + blockIdx.x;
    const int row = la.rows[r];
    const int a0 = la.arpt[row];
    const int L = la.arpt[row + 1] - a0;
    const int64_t* ipr = la.ipA + a0;
    const int H = ceil_log2i(L);
    const int pb = la.pass_bits;
    const int J = H == 0 ? 1 : (H + pb - 1) / pb;
    const WsBase wsb{la.ws_base[2 * (int64_t)r], la.ws_base[2 * (int64_t)r + 1]};
    const int64_t slb = la.slot_base[r];
    const int64_t final_off = la.final_off ? la.final_off[row] : ipr[0];
    if (la.frontier_cap == 0 && (1 << pb) <= PL_NT) {
        plan_no_frontier(la, row, a0, ipr, L, H, J, wsb, slb, final_off, err);
        if (threadIdx.x == 0) atomicMax((unsigned long long*)&la.counters[CNT_MAXP], (unsigned long long)J);
        return;
    }

    for (int p = threadIdx.x; p < L; p += PL_NT) {
        const int hf = frontier_height(ipr, L, H, p, cap);
        const int s = p & ~((1 << hf) - 1);
        if (s != p) continue;
        const NodeK fk = node_kind(ipr, L, s, hf);
        if (hf == H) {  // the whole row is one frontier node
            if (fk.kind == SRC_STORED) {
                const unsigned long long li = atomicAdd((unsigned long long*)&la.counters[CNT_LEAF], 1ull);
                atomicAdd((unsigned long long*)&la.counters[CNT_LEAF_IP], (unsigned long long)fk.ip);
                if ((int64_t)li >= la.leaves_cap) { atomicOr(err, (int)DERR_CAPACITY); continue; }
                LeafTask lt;
                lt.row = row;
                lt.ab = a0;
                lt.ae = a0 + L;
                lt.buf = 0;
                lt.final_dst = 1;
                lt.pad = 0;
                lt.dst_off = final_off;
                lt.slot = -1;
                la.leaves[li] = lt;
            } else if (fk.kind == SRC_VIRTUAL) {
                emit_group(la, row, a0, ipr, L, H, 1, 0, H, 0, cap, wsb, slb, final_off, err);
            }
            continue;
        }
        const int jf = hf / pb + 1;  // pass that consumes this frontier node
        if (fk.kind == SRC_STORED) {
            const unsigned long long li = atomicAdd((unsigned long long*)&la.counters[CNT_LEAF], 1ull);
            atomicAdd((unsigned long long*)&la.counters[CNT_LEAF_IP], (unsigned long long)fk.ip);
            if ((int64_t)li >= la.leaves_cap) { atomicOr(err, (int)DERR_CAPACITY); continue; }
            LeafTask lt;
            lt.row = row;
            lt.ab = a0 + s;
            lt.ae = a0 + node_end(L, s, hf);
            lt.buf = (jf - 1) & 1;
            lt.final_dst = 0;
            lt.pad = 0;
            lt.dst_off = wsb.at(lt.buf) + (ipr[s] - ipr[0]);
            lt.slot = slb + heap_idx(H, s, hf);
            la.leaves[li] = lt;
        }
        // pass nodes whose leftmost frontier node starts at s
        for (int j = jf; j <= J; ++j) {
            const int hj = pass_height(j, H, pb);
            if (s & ((1 << hj) - 1)) break;
            emit_group(la, row, a0, ipr, L, H, j, s, hj, pass_height(j - 1, H, pb), cap, wsb, slb, final_off, err);
        }
    }
    if (threadIdx.x == 0) atomicMax((unsigned long long*)&la.counters[CNT_MAXP], (unsigned long long)J);
}

// ---------------------------------------------------------------- group merge
struct WsView {
    int* col0;
    int* col1;
    double* val0;
    double* val1;
};

template <int GNT, int GCAP>
struct GroupSmem {
    MergeSmem<GNT, GCAP, GROUP_KMAX> m;
    double scale[GROUP_KMAX];
    int64_t off[GROUP_KMAX];
    int64_t len[GROUP_KMAX];
    int64_t cur[GROUP_KMAX];
    int kind[GROUP_KMAX];
    int buf[GROUP_KMAX];
    int lid[GROUP_KMAX];
    int w[GROUP_KMAX];
    int woff[GROUP_KMAX + 1];
    int take[GROUP_KMAX];
    int orig[GROUP_KMAX];
    int nlists;
    int done;
};

// Window sizes of the next round (wavefront 0 computes, all threads read):
// proportional to what each child has left, all of it when everything fits.
template <int GNT, int GCAP>
__device__ __forceinline__ void group_windows(GroupSmem<GNT, GCAP>& g, int k) {
    const int t = threadIdx.x;
    if (t < BRM_WAVE) {
        const int64_t rem = t < k ? g.len[t] - g.cur[t] : 0;
        const int64_t R = wave_sum(rem);
        int64_t wv;
        if (R <= GCAP) wv = rem;
        else wv = rem ? max((int64_t)1, (int64_t)(GCAP - GROUP_KMAX) * rem / R) : 0;
        if (wv > rem) wv = rem;
        const int wi = (int)wv;
        const int inc = wave_incl_scan(wi);
        if (t < k) {
            g.w[t] = wi;
            g.woff[t] = inc - wi;
        }
        if (t == BRM_WAVE - 1) {
            g.woff[k] = inc;
            g.done = (R == 0);
        }
    }
}

// Issues the global loads of the current windows into registers.
template <int GNT, int GCAP, int ITEMS>
__device__ __forceinline__ void group_fetch(GroupSmem<GNT, GCAP>& g, int k, const RowArgs& a, const WsView& ws,
                                            int (&rc)[ITEMS], double (&rv)[ITEMS], double (&rs)[ITEMS],
                                            unsigned& vmask) {
    const int W = g.woff[k];
    int p = 0;
    vmask = 0u;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int e = threadIdx.x + i * GNT;
        if (e < W) {
            p = gallop_idx(g.woff, p, k, e);
            const int64_t x = g.off[p] + g.cur[p] + (e - g.woff[p]);
            const bool virt = g.kind[p] == SRC_VIRTUAL;
            const int* cbase = virt ? a.bcol : (g.buf[p] ? ws.col1 : ws.col0);
            const double* vbase = virt ? a.bval : (g.buf[p] ? ws.val1 : ws.val0);
            rc[i] = cbase[x];
            rv[i] = vbase[x];
            rs[i] = g.scale[p];
            if (virt) vmask |= 1u << i;  // virtual: the product is rounded when stored
        }
    }
}

template <int GNT, int GCAP>
__global__ __launch_bounds__(GNT) void k_group(const GroupTask* __restrict__ tasks,
                                               const Child* __restrict__ children, RowArgs a, WsView ws,
                                               int64_t* __restrict__ slots, int64_t* counters, OutArgs o) {
    constexpr int ITEMS = GCAP / GNT;
    __shared__ GroupSmem<GNT, GCAP> g;
    const int t = threadIdx.x;
    const int lane = lane_id();
    const GroupTask T = tasks[blockIdx.x];
    const int k = T.nchild;
    if (t < k) {
        const Child c = children[T.cbase + t];
        g.kind[t] = c.src.kind;
        g.buf[t] = c.src.buf;
        g.off[t] = c.src.off;
        g.scale[t] = c.src.scale;
        g.lid[t] = c.lid;
        g.len[t] = c.src.kind == SRC_VIRTUAL ? c.src.len : slots[c.src.slot];
        g.cur[t] = 0;
    }
    __syncthreads();

    int* dc;
    double* dv;
    int64_t dcap;
    if (T.final_dst) {
        dc = o.ccol + T.dst_off;
        dv = o.cval + T.dst_off;
        dcap = o.upper ? INT64_MAX : o.crpt[T.row + 1] - o.crpt[T.row];
    } else {
        dc = (T.dst_buf ? ws.col1 : ws.col0) + T.dst_off;
        dv = (T.dst_buf ? ws.val1 : ws.val0) + T.dst_off;
        dcap = INT64_MAX;
    }

    int64_t in_total = 0;
    if (t < BRM_WAVE) {
        const int64_t l = t < k ? g.len[t] : 0;
        in_total = wave_sum(l);
    }
    int64_t out_n = 0;
    int rounds = 0;
    static_assert(ITEMS <= 32, "window items per thread");
    int rc[ITEMS];
    double rv[ITEMS], rs[ITEMS];
    unsigned vmask = 0u;

    group_windows<GNT, GCAP>(g, k);
    __syncthreads();
    bool done = g.done;
    if (!done) group_fetch<GNT, GCAP, ITEMS>(g, k, a, ws, rc, rv, rs, vmask);
    while (!done) {
        rounds++;
        // ---- the fetched windows into buffer 1 (products rounded here)
        {
            const int W = g.woff[k];
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const int e = t + i * GNT;
                if (e < W) {
                    g.m.col[1][e] = rc[i];
                    g.m.val[1][e] = ((vmask >> i) & 1u) ? fmul(rs[i], rv[i]) : rv[i];
                }
            }
        }
        __syncthreads();
        // ---- merge bound: smallest last column of a partially loaded child;
        //      every child contributes its window prefix up to that bound
        if (t < BRM_WAVE) {
            int lastc = INT_MAX;
            if (t < k && g.w[t] < g.len[t] - g.cur[t]) lastc = g.m.col[1][g.woff[t] + g.w[t] - 1];
#pragma unroll
            for (int d = BRM_WAVE / 2; d > 0; d >>= 1) lastc = min(lastc, __shfl_xor(lastc, d, BRM_WAVE));
            int take = 0;
            if (t < k) {
                const int* wc = g.m.col[1] + g.woff[t];
                int lo = 0, hi = g.w[t];
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (wc[mid] <= lastc) lo = mid + 1; else hi = mid;
                }
                take = lo;
            }
            const uint64_t nz = __ballot(take > 0);
            const int idx = __popcll(nz & ((1ull << lane) - 1ull));
            const int ex = wave_incl_scan(take) - take;
            if (take > 0) {
                g.m.lid[0][idx] = g.lid[t];
                g.m.loff[0][idx] = ex;
                g.orig[idx] = t;
                g.take[idx] = g.woff[t];  // window start of the idx-th taken list
            }
            if (t == BRM_WAVE - 1) {
                g.nlists = __popcll(nz);
                g.m.loff[0][__popcll(nz)] = ex + take;
            }
            if (t < k) g.cur[t] += take;  // consumed: the next windows start after it
        }
        __syncthreads();
        // ---- gather the taken prefixes into buffer 0, contiguous, in child order
        const int nl = g.nlists;
        const int S = g.m.loff[0][nl];
        {
            int q = 0;
            for (int e = t; e < S; e += GNT) {
                q = gallop_idx(g.m.loff[0], q, nl, e);
                const int src = g.take[q] + (e - g.m.loff[0][q]);
                g.m.col[0][e] = g.m.col[1][src];
                g.m.val[0][e] = g.m.val[1][src];
            }
        }
        // ---- next round's windows; their loads overlap this round's merge
        group_windows<GNT, GCAP>(g, k);
        __syncthreads();
        done = g.done;
        if (!done) group_fetch<GNT, GCAP, ITEMS>(g, k, a, ws, rc, rv, rs, vmask);
        int cb;
        const int len = merge_levels<GNT, GCAP, GROUP_KMAX>(g.m, nl, &cb);
        // ---- stream out
        for (int e = t; e < len; e += GNT) {
            const int64_t o_ = out_n + e;
            if (o_ < dcap) {
                dc[o_] = g.m.col[cb][e];
                dv[o_] = g.m.val[cb][e];
            }
        }
        out_n += len;
        __syncthreads();
    }

    if (t == 0) {
        if (T.final_dst) {
            if (o.upper) o.rowlen[T.row] = out_n;
            else if (out_n != dcap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
        } else {
            slots[T.slot] = out_n;
        }
        atomicAdd((unsigned long long*)&counters[CNT_GRP_IN], (unsigned long long)in_total);
        atomicAdd((unsigned long long*)&counters[CNT_GRP_OUT], (unsigned long long)out_n);
        atomicAdd((unsigned long long*)&counters[CNT_ROUNDS], (unsigned long long)rounds);
    }
}

}  // namespace

hipError_t launch_plan(const LongArgs& la, int* err, hipStream_t s) {
    if (la.nrows <= 0) return hipSuccess;
    {
        ProfScope ps_("k_plan", s);
        const int64_t ch = launch_chunk(PL_NT);
        for (int64_t b = 0; b < la.nrows; b += ch)
            k_plan<<<(unsigned)std::min<int64_t>(ch, la.nrows - b), PL_NT, 0, s>>>(la, (int)b, err);
    }
    return hipGetLastError();
}

hipError_t launch_group(const GroupTask* tasks, int64_t n, const Child* children, const RowArgs& a,
                        int* const ws_col[2], double* const ws_val[2], int64_t* slots, int64_t* counters,
                        const OutArgs& o, hipStream_t s, int group_cfg, const int* order) {
    if (n <= 0) return hipSuccess;
    if (group_cfg == GROUP_CFG_BITMAP)
        return launch_bgroup(tasks, n, children, a, ws_col, ws_val, slots, counters, o, s, order);
    if (order) return hipErrorInvalidValue;
    WsView w{ws_col[0], ws_col[1], ws_val[0], ws_val[1]};
    ProfScope ps_("k_group", s);
    const int64_t ch = launch_chunk(256);
    for (int64_t b = 0; b < n; b += ch) {
        const unsigned c = (unsigned)std::min<int64_t>(ch, n - b);
        const GroupTask* tb = tasks + b;
        switch (group_cfg) {
            case 1: k_group<128, 2048><<<c, 128, 0, s>>>(tb, children, a, w, slots, counters, o); break;
            case 2: k_group<64, 1024><<<c, 64, 0, s>>>(tb, children, a, w, slots, counters, o); break;
            case 3: k_group<256, 4096><<<c, 256, 0, s>>>(tb, children, a, w, slots, counters, o); break;
            default: k_group<256, 2048><<<c, 256, 0, s>>>(tb, children, a, w, slots, counters, o); break;
        }
    }
    return hipGetLastError();
}

namespace {
__global__ void k_iota(int* __restrict__ p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = i;
}
}  // namespace

namespace {
__global__ void k_row_ip_keys(const int* __restrict__ rows, int n, const int* __restrict__ arpt,
                              const int64_t* __restrict__ ipA, uint64_t* __restrict__ keys) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int r = rows[i];
    keys[i] = (uint64_t)(ipA[arpt[r + 1]] - ipA[arpt[r]]);
}
}  // namespace

size_t order_rows_tmp_bytes(int n) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                     (const int*)nullptr, (int*)nullptr, n, 0, 64) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return bytes;
}

hipError_t order_rows(int* rows, int n, const int* arpt, const int64_t* ipA, uint64_t* keys, uint64_t* keys_out,
                      int* rows_out, void* tmp, size_t tmp_bytes, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    k_row_ip_keys<<<(n + 255) / 256, 256, 0, s>>>(rows, n, arpt, ipA, keys);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    size_t bytes = tmp_bytes;
    if (hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(tmp, bytes, keys, keys_out, rows, rows_out, n, 0,
                                                                    64, s);
        e != hipSuccess)
        return e;
    return hipMemcpyAsync(rows, rows_out, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s);
}

size_t order_tasks_tmp_bytes(int n) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                     (const int*)nullptr, (int*)nullptr, n, 0, 32) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return bytes;
}

hipError_t order_tasks(const uint32_t* keys, int n, uint32_t* keys_out, int* idx_in, int* idx_out, void* tmp,
                       size_t tmp_bytes, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_iota<<<(n + 255) / 256, 256, 0, s>>>(idx_in, n);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    size_t bytes = tmp_bytes;
    return hipcub::DeviceRadixSort::SortPairsDescending(tmp, bytes, keys, keys_out, idx_in, idx_out, n, 0, 32, s);
}

}  // namespace brm
