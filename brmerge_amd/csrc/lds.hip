// LDS tier: one workgroup merges one row (or one aligned subtree of a long
// row's lists) entirely in LDS, and the symbolic counters for medium and
// long rows.
//
// Tree merge (BRMerge-SpGEMM/inc/spgemm.h:610-707 restated for a workgroup):
//   EXPAND   every non-empty list p (B row acol[p] scaled by aval[p]) is laid
//            out contiguously in LDS buffer 0 (coalesced B gathers);
//   MERGE    level by level, lists whose node ids share a parent
//            (id >> 1 equal) form a pair, others a single that is carried;
//            all pairs of a level are merged at once with merge-path
//            partitioning of the level's raw element range over the
//            workgroup's threads (a count pass, a block scan, a write pass),
//            from one LDS buffer into the other (ping-pong);
//   OUTPUT   the last list is streamed to C / C-bar / the workspace.
// Empty lists are never materialised but their positions are kept in the
// node ids, so the pairing -- and therefore the bracketing of every sum --
// is that of the reference's aligned tree.
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"
#include "lds_merge.h"

namespace brm {
namespace {

// Merge the non-empty lists of A entries [ab, ae) of one row.  Positions are
// taken relative to ab, which must be aligned to the subtree size.  The
// result (length returned, uniform) is written to dcol/dval.
template <int NT, int CAP, int LMAX>
__device__ int64_t lds_tree_merge(MergeSmem<NT, CAP, LMAX>& sm, const RowArgs& a, int ab, int ae,
                                  int* __restrict__ dcol, double* __restrict__ dval, int64_t dcap, int* err) {
    const int t = threadIdx.x;
    const int64_t ibase = a.ipA[ab];
    const int S = (int)(a.ipA[ae] - ibase);

    // ---- list descriptors of the non-empty lists: node id, offset, and (in
    //      the second ping-pong buffers, unused until the first merge level)
    //      the B row start and the A value, so that the expansion below issues
    //      independent loads only
    int* bst = sm.lid[1];
    double* sc = sm.val[1];
    int n = 0;
    for (int cs = ab; cs < ae; cs += NT) {
        const int ai = cs + t;
        int len = 0;
        if (ai < ae) len = (int)(a.ipA[ai + 1] - a.ipA[ai]);
        const int flag = len > 0;
        int tot;
        const int ex = block_excl_scan<NT>(flag, sm.scan, &tot);
        if (flag && n + ex < LMAX) {
            sm.lid[0][n + ex] = ai - ab;
            sm.loff[0][n + ex] = (int)(a.ipA[ai] - ibase);
            bst[n + ex] = a.brpt[a.acol[ai]];
            sc[n + ex] = a.aval[ai];
        }
        n += tot;
    }
    if (n > LMAX || S > CAP) {  // planner guarantees otherwise
        if (t == 0) atomicOr(err, (int)DERR_CAPACITY);
        return 0;
    }
    if (t == 0) sm.loff[0][n] = S;
    __syncthreads();

    // ---- EXPAND (spgemm.h:618-624): products rounded before any add; all of
    //      a thread's loads are issued before its first LDS store
    {
        constexpr int ITEMS = CAP / NT;
        int rc[ITEMS];
        double rv[ITEMS], rs[ITEMS];
        int k = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = t + i * NT;
            if (e < S) {
                k = gallop_idx(sm.loff[0], k, n, e);
                const int bi = bst[k] + (e - sm.loff[0][k]);
                rc[i] = a.bcol[bi];
                rv[i] = a.bval[bi];
                rs[i] = sc[k];
            }
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = t + i * NT;
            if (e < S) {
                sm.col[0][e] = rc[i];
                sm.val[0][e] = fmul(rs[i], rv[i]);
            }
        }
    }
    __syncthreads();

    int cur = 0;
    int len = merge_levels<NT, CAP, LMAX>(sm, n, &cur);

    // ---- OUTPUT
    const int wlen = (int)min((int64_t)len, dcap);  // never write past the destination
    for (int e = t; e < wlen; e += NT) {
        dcol[e] = sm.col[cur][e];
        dval[e] = sm.val[cur][e];
    }
    return len;
}

template <int NT, int CAP, int LMAX>
__global__ __launch_bounds__(NT) void k_medium(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o) {
    extern __shared__ __align__(16) unsigned char smraw[];
    auto& sm = *reinterpret_cast<MergeSmem<NT, CAP, LMAX>*>(smraw);
    const int row = rows[blockIdx.x];
    const int ab = a.arpt[row], ae = a.arpt[row + 1];
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        const int64_t cap = o.crpt[row + 1] - cb;
        const int64_t len = lds_tree_merge<NT, CAP, LMAX>(sm, a, ab, ae, o.ccol + cb, o.cval + cb, cap, o.err);
        if (threadIdx.x == 0 && len != cap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
    } else {
        const int64_t cb = a.ipA[ab];
        const int64_t len =
            lds_tree_merge<NT, CAP, LMAX>(sm, a, ab, ae, o.ccol + cb, o.cval + cb, a.ipA[ae] - cb, o.err);
        if (threadIdx.x == 0) o.rowlen[row] = len;
    }
}

struct WsPtrs {
    int* col0;
    int* col1;
    double* val0;
    double* val1;
};

template <int NT, int CAP, int LMAX>
__global__ __launch_bounds__(NT) void k_leaf(const LeafTask* __restrict__ tasks, RowArgs a, WsPtrs ws,
                                             int64_t* __restrict__ slots, int64_t* counters, OutArgs o) {
    extern __shared__ __align__(16) unsigned char smraw[];
    auto& sm = *reinterpret_cast<MergeSmem<NT, CAP, LMAX>*>(smraw);
    const LeafTask tk = tasks[blockIdx.x];
    int* dc;
    double* dv;
    int64_t cap;
    if (tk.final_dst) {
        dc = o.ccol + tk.dst_off;
        dv = o.cval + tk.dst_off;
        cap = o.upper ? a.ipA[tk.ae] - a.ipA[tk.ab] : o.crpt[tk.row + 1] - o.crpt[tk.row];
    } else {
        dc = (tk.buf ? ws.col1 : ws.col0) + tk.dst_off;
        dv = (tk.buf ? ws.val1 : ws.val0) + tk.dst_off;
        cap = a.ipA[tk.ae] - a.ipA[tk.ab];
    }
    const int64_t len = lds_tree_merge<NT, CAP, LMAX>(sm, a, tk.ab, tk.ae, dc, dv, cap, o.err);
    if (threadIdx.x == 0) {
        if (tk.final_dst) {
            if (o.upper) o.rowlen[tk.row] = len;
            else if (len != cap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
        } else {
            slots[tk.slot] = len;
        }
        atomicAdd((unsigned long long*)&counters[CNT_LEAF_OUT], (unsigned long long)len);
    }
}

// ---------------------------------------------------------------- symbolic
// Medium rows: distinct columns counted with an LDS open-addressing table
// (the reference uses a per-thread linear-probing table of upper_log(ip)
// slots, spgemm.h:528-556; any table gives the same count).
template <int HS_NT, int HS_CAP, int HS_LMAX>
__global__ __launch_bounds__(HS_NT) void k_sym_hash(const int* __restrict__ rows, int nrows, RowArgs a,
                                                    int64_t* __restrict__ nnzrow) {
    constexpr int ITEMS = 8;  // loads in flight per thread
    __shared__ int table[HS_CAP];
    __shared__ int loff[HS_LMAX + 1];
    __shared__ int bst[HS_LMAX];
    __shared__ int sm[HS_NT / BRM_WAVE + 1];
    const int t = threadIdx.x;
    const int row = rows[blockIdx.x];
    const int ab = a.arpt[row];
    const int L = a.arpt[row + 1] - ab;
    const int64_t ib = a.ipA[ab];
    const int S = (int)(a.ipA[ab + L] - ib);
    int hs = 64;
    while (hs < 2 * S) hs <<= 1;
    for (int k = t; k < hs; k += HS_NT) table[k] = -1;
    for (int k = t; k <= L; k += HS_NT) {
        loff[k] = (int)(a.ipA[ab + k] - ib);
        if (k < L) bst[k] = a.brpt[a.acol[ab + k]];
    }
    __syncthreads();
    int cnt = 0;
    int hint = 0;
    for (int e0 = 0; e0 < S; e0 += ITEMS * HS_NT) {
        int key[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int e = e0 + t + i * HS_NT;
            key[i] = -1;
            if (e < S) {
                hint = gallop_idx(loff, hint, L, e);
                key[i] = a.bcol[bst[hint] + (e - loff[hint])];
            }
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (key[i] < 0) continue;
            unsigned h = ((unsigned)key[i] * 2654435761u) & (unsigned)(hs - 1);
            for (;;) {
                const int old = atomicCAS(&table[h], -1, key[i]);
                if (old == -1) { cnt++; break; }
                if (old == key[i]) break;
                h = (h + 1) & (unsigned)(hs - 1);
            }
        }
    }
    const int tot = block_sum<HS_NT>(cnt, sm);
    if (t == 0) nnzrow[row] = tot;
}

// Long rows: distinct columns counted with an LDS bitmap of up to 2^20
// columns per pass (128 KB); wider matrices take several column passes.
// Lists are taken in chunks of BM_CHUNK: their product offsets and B row
// starts go to LDS first, so a product's column is one global load, and
// each thread handles BM_U consecutive products per step with their loads
// issued together.
constexpr int BM_NT = 1024;
constexpr int BM_BITS = 1 << 20;
constexpr int BM_WORDS = BM_BITS / 32;
constexpr int BM_CHUNK = 2048;
constexpr int BM_U = 8;

// A chunk's product offsets are kept in int: chunks are cut at BM_CHUNK lists
// and at pcap products (at most 2^30, so offsets and the loop counters below
// stay inside int).  One list is always taken: a B row has at most
// INT32_MAX entries, and a list longer than pcap is its own chunk.
__device__ __forceinline__ int sym_chunk_lists(const int64_t* __restrict__ ipA, int cs, int ae, int64_t pcap) {
    int nl = min(BM_CHUNK, ae - cs);
    const int64_t cb = ipA[cs];
    if (ipA[cs + nl] - cb <= pcap) return nl;
    int lo = 1, hi = nl;  // ipA[cs + hi] - cb > pcap
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ipA[cs + mid] - cb <= pcap) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(BM_NT) void k_sym_bitmap(const int* __restrict__ rows, int nrows, RowArgs a,
                                                      int64_t* __restrict__ nnzrow, int64_t pcap) {
    extern __shared__ __align__(16) unsigned char smraw[];
    unsigned* bm = reinterpret_cast<unsigned*>(smraw);
    int* loff = reinterpret_cast<int*>(smraw + sizeof(unsigned) * BM_WORDS);
    int* bst = loff + BM_CHUNK + 8;
    int64_t* red = reinterpret_cast<int64_t*>(bst + BM_CHUNK + 8);
    const int t = threadIdx.x;
    const int row = rows[blockIdx.x];
    const int ab = a.arpt[row], ae = a.arpt[row + 1];
    int64_t total = 0;
    for (int64_t c0 = 0; c0 < a.N; c0 += BM_BITS) {
        const int64_t span = min((int64_t)BM_BITS, (int64_t)a.N - c0);
        const int words = (int)((span + 31) / 32);
        const int c0i = (int)c0;
        const unsigned span32 = (unsigned)span;
        for (int w = t; w < words; w += BM_NT) bm[w] = 0u;
        __syncthreads();
        for (int cs = ab, nl = 0; cs < ae; cs += nl) {
            nl = sym_chunk_lists(a.ipA, cs, ae, pcap);
            const int64_t cb = a.ipA[cs];
            for (int k = t; k <= nl; k += BM_NT) {
                loff[k] = (int)(a.ipA[cs + k] - cb);
                if (k < nl) bst[k] = a.brpt[a.acol[cs + k]];
            }
            __syncthreads();
            const int Sc = loff[nl];
            int hint = 0;
            for (int e0 = t * BM_U; e0 < Sc; e0 += BM_NT * BM_U) {
                hint = gallop_idx(loff, hint, nl, e0);
                int col[BM_U];
                int h = hint;
                if (e0 + BM_U <= loff[h + 1]) {
                    // all BM_U products in list h: consecutive entries of one
                    // B row, loaded without a list search per product
                    const int* p = a.bcol + bst[h] + (e0 - loff[h]);
#pragma unroll
                    for (int u = 0; u < BM_U; ++u) col[u] = p[u];
                } else {
#pragma unroll
                    for (int u = 0; u < BM_U; ++u) {
                        const int e = e0 + u;
                        col[u] = -1;
                        if (e < Sc) {
                            while (loff[h + 1] <= e) ++h;  // next non-empty list
                            col[u] = a.bcol[bst[h] + (e - loff[h])];
                        }
                    }
                }
                // 32-bit offsets: columns and c0 are below 2^31; a column
                // below c0 (or -1, no product) wraps past the span
#pragma unroll
                for (int u = 0; u < BM_U; ++u) {
                    const unsigned rel = (unsigned)(col[u] - c0i);
                    if (rel < span32) atomicOr(&bm[rel >> 5], 1u << (rel & 31));
                }
            }
            __syncthreads();
        }
        int64_t cnt = 0;
        for (int w = t; w < words; w += BM_NT) cnt += __popc(bm[w]);
        total += block_sum<BM_NT>(cnt, red);
        __syncthreads();
    }
    if (t == 0) nnzrow[row] = total;
}

constexpr size_t bitmap_smem() {
    return sizeof(unsigned) * BM_WORDS + 2 * sizeof(int) * (BM_CHUNK + 8) + sizeof(int64_t) * (BM_NT / BRM_WAVE + 2);
}
static_assert(bitmap_smem() <= 160 * 1024, "LDS of one CU");

template <int NT, int CAP, int LMAX>
hipError_t medium_numeric_t(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                            const char* name = "k_medium") {
    using SM = MergeSmem<NT, CAP, LMAX>;
    auto kern = k_medium<NT, CAP, LMAX>;
    static SmemLimit lim;
    if (hipError_t e = lim.raise((const void*)kern, (int)sizeof(SM)); e != hipSuccess) return e;
    {
        ProfScope ps_(name, s);
        const int64_t ch = launch_chunk(NT);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            kern<<<c, NT, sizeof(SM), s>>>(rows + b, c, a, o);
        }
    }
    return hipGetLastError();
}

template <int NT, int CAP, int LMAX>
hipError_t leaf_t(const LeafTask* tasks, int64_t n, const RowArgs& a, const WsPtrs& ws, int64_t* slots,
                  int64_t* counters, const OutArgs& o, hipStream_t s) {
    using SM = MergeSmem<NT, CAP, LMAX>;
    auto kern = k_leaf<NT, CAP, LMAX>;
    static SmemLimit lim;
    if (hipError_t e = lim.raise((const void*)kern, (int)sizeof(SM)); e != hipSuccess) return e;
    {
        ProfScope ps_("k_leaf", s);
        const int64_t ch = launch_chunk(NT);
        for (int64_t b = 0; b < n; b += ch)
            kern<<<(unsigned)std::min<int64_t>(ch, n - b), NT, sizeof(SM), s>>>(tasks + b, a, ws, slots, counters, o);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_medium_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                                 int lds_cfg) {
    if (n <= 0) return hipSuccess;
    switch (lds_cfg) {
        case 1: return medium_numeric_t<256, 2048, 256>(rows, n, a, o, s);
        case 2: return medium_numeric_t<128, 2048, 256>(rows, n, a, o, s);
        default: return medium_numeric_t<256, 4096, 512>(rows, n, a, o, s);
    }
}

hipError_t launch_leaf(const LeafTask* tasks, int64_t n, const RowArgs& a, int* const ws_col[2],
                       double* const ws_val[2], int64_t* slots, int64_t* counters, const OutArgs& o,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    WsPtrs ws{ws_col[0], ws_col[1], ws_val[0], ws_val[1]};
    switch (params().lds_cfg) {
        case 1: return leaf_t<256, 2048, 256>(tasks, n, a, ws, slots, counters, o, s);
        case 2: return leaf_t<128, 2048, 256>(tasks, n, a, ws, slots, counters, o, s);
        default: return leaf_t<256, 4096, 512>(tasks, n, a, ws, slots, counters, o, s);
    }
}

hipError_t launch_medium_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_sym_hash", s);
        const int64_t ch = launch_chunk(256);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            k_sym_hash<256, 2 * MED_CAP, MED_LMAX><<<c, 256, 0, s>>>(rows + b, c, a, nnzrow);
        }
    }
    return hipGetLastError();
}

// Wavefront tier: one 64-thread workgroup per row, so several rows share a
// CU and no row waits on another's barriers.
hipError_t launch_wmed_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s,
                             int wmed_cap) {
    if (n <= 0) return hipSuccess;
    if (params().wmed_count == 1 && rmed_count_fits(a.N)) return launch_rmed_count(rows, n, a, nnzrow, s, wmed_cap);
    {
        ProfScope ps_("k_sym_hash_w", s);
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            // table of twice the row's largest product count (smaller LDS, more waves)
            if (wmed_cap <= 256) k_sym_hash<BRM_WAVE, 512, WMED_LMAX><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
            else if (wmed_cap <= 512) k_sym_hash<BRM_WAVE, 1024, WMED_LMAX><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
            else if (wmed_cap <= WMED_CAP) k_sym_hash<BRM_WAVE, 2 * WMED_CAP, WMED_LMAX><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
            else k_sym_hash<BRM_WAVE, 2 * WMED_CAP_BITMAP, WMED_LMAX><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
        }
    }
    return hipGetLastError();
}

hipError_t launch_wmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                               int wmed_cap) {
    if (n <= 0) return hipSuccess;
    if (params().wmed_engine == 1) return launch_rmed_numeric(rows, n, a, o, s, wmed_cap);
    if (wmed_cap > WMED_CAP) return hipErrorInvalidValue;  // binned only for the register tier
    if (wmed_cap <= 256) return medium_numeric_t<BRM_WAVE, 256, WMED_LMAX>(rows, n, a, o, s, "k_medium_w");
    if (wmed_cap <= 512) return medium_numeric_t<BRM_WAVE, 512, WMED_LMAX>(rows, n, a, o, s, "k_medium_w");
    return medium_numeric_t<BRM_WAVE, WMED_CAP, WMED_LMAX>(rows, n, a, o, s, "k_medium_w");
}

hipError_t launch_hash_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_sym_hash_h", s);
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            if (cap <= HASH_SMALL)
                k_sym_hash<BRM_WAVE, 2 * HASH_SMALL, HASH_SMALL><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
            else
                k_sym_hash<BRM_WAVE, 2 * HASH_CAP, HASH_CAP><<<c, BRM_WAVE, 0, s>>>(rows + b, c, a, nnzrow);
        }
    }
    return hipGetLastError();
}

hipError_t launch_long_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    static SmemLimit lim;
    if (hipError_t e = lim.raise((const void*)k_sym_bitmap, (int)bitmap_smem()); e != hipSuccess) return e;
    {
        ProfScope ps_("k_sym_bitmap", s);
        const int64_t ch = launch_chunk(BM_NT);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            k_sym_bitmap<<<c, BM_NT, bitmap_smem(), s>>>(rows + b, c, a, nnzrow, params().sym_chunk_products);
        }
    }
    return hipGetLastError();
}

}  // namespace brm
