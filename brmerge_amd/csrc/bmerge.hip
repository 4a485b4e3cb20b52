// Bitmap-rank group merge for long rows (group_cfg 4).
//
// One workgroup materialises one tree node above the lists of a long row
// from up to BGROUP_KMAX child lists (B rows read in place, or nodes of the previous
// pass), in rounds over a streaming window, like the merge-path engine, but
// without merge levels:
//   1. every child loads a window proportional to what it has left; then
//      (BG_TOPUPS times) the children whose windows stop short of a target
//      column (the lower edge of the median log2-extent among partially
//      loaded children) share the free window space as top-ups; all loaded
//      elements with a column <= bound (the smallest last loaded column of
//      a partially loaded child, and at most c_lo + 2^16 - 1) are taken --
//      every contribution to those columns is in the round;
//   2. the taken columns set bits of an LDS bitmap over [c_lo, c_lo + 2^16);
//      a popcount prefix over the bitmap words ranks every distinct column,
//      which is its output position in the round;
//   3. contributions are bucketed by output position (counting sort), each
//      bucket is ordered by child, i.e. by list position, and summed as the
//      aligned binary tree over list positions sums them
//      (BRMerge-SpGEMM/inc/spgemm.h:639-707: siblings added left + right,
//      a node with one present child passes it through).
// The work per element does not depend on the number of children.  The
// fan-in (BGROUP_KMAX = 2^BPASS_BITS) trades the take rate against the
// number of HBM passes: the bound is the smallest end among the partially
// loaded children, so the longer each child's share of the window, the more
// of the window is taken per round, while every pass streams the row's
// nodes through HBM once more.  With 2^6 children a round of R-MAT-20 takes
// 1,120 of its 2,048 elements (1,156 with 2^5), and 9% fewer elements pass
// through the engine (rows of 33 to 64 lists take one pass, of 1,025 to
// 4,096 two); 2^7 children do not fit three workgroups per CU.
#include <limits.h>
#include <stddef.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int BG_NT = 256;
constexpr int BG_W = 2048;                    // window elements per round
constexpr int BG_IPT = BG_W / BG_NT;          // window elements per thread
constexpr int BG_SPAN_LOG2 = 16;
constexpr int64_t BG_SPAN = 1 << BG_SPAN_LOG2;  // columns per round
constexpr int BG_WORDS = (int)(BG_SPAN / 64);
constexpr int BG_WPT = BG_WORDS / BG_NT;      // bitmap words per thread
constexpr int BG_KMAX = BGROUP_KMAX;
constexpr int BG_NW = BG_NT / BRM_WAVE;
constexpr int BG_NSLOT = 8;    // rotating reduction scratch slots
constexpr int BG_TOPUPS = 1;   // top-up iterations per round
// Outputs with more contributions than this are summed by a whole wavefront
// (dense over the list ids) instead of by one thread.
constexpr int BG_THREAD_M = 8;
constexpr int BG_BIGMAX = BG_W / (BG_THREAD_M + 1) + 1;
constexpr int BG_MIDMAX = BG_W / 3 + 1;  // outputs with 3 .. BG_THREAD_M contributions
static_assert(BG_KMAX <= BG_NT, "one child per thread in the window logic");
static_assert(BG_W <= 65535, "u16 ranks and contribution indices");

struct BgSmem {
    double val[BG_W];
    int col[BG_W];
    unsigned long long bm[BG_WORDS];
    unsigned short cnt[BG_W];  // bucket counts, then offsets, then ends (atomics: cnt_add)
    unsigned short wrank[BG_WORDS];
    unsigned short contrib[BG_W];
    unsigned char elid[BG_W];       // list id (node id) of the element's child
    int64_t off[BG_KMAX];
    int64_t len[BG_KMAX];
    int64_t cur[BG_KMAX];
    double scale[BG_KMAX];
    int w[BG_KMAX];         // elements of the child loaded this round
    int woff[BG_KMAX + 1];  // proportional windows: offsets in the buffer
    unsigned long long wstart[BG_W / 64];  // bit woff[j] set for every child with a window
    unsigned char wchild[BG_KMAX];         // the children with windows, in order
    int lastc[BG_KMAX];     // last loaded column of the child
    unsigned long long xstart[BG_W / 64];  // top-up runs: bit at each run's start (relative to the fill level)
    unsigned char xchild[BG_KMAX];         // the children with top-up runs, in order
    int64_t xbase[BG_KMAX];  // top-up element e of the child is at xbase + e
    int hist[BG_TOPUPS][32];  // per top-up: partial children by log2 extent
    int64_t fbase[BG_KMAX];  // off + cur - woff: element e of the child's window is at fbase + e
    unsigned char src[BG_KMAX];  // SRC_WS0 / SRC_WS1 / SRC_B
    unsigned char lid[BG_KMAX];
    unsigned short big[BG_BIGMAX];  // outputs left to the wavefront sums
    unsigned short mid[BG_MIDMAX];  // outputs left to the second thread pass
    int nbig;
    int nmid;
    int bnext;  // next big bucket to claim
    int64_t red64[BG_NSLOT][BG_NW];
    int red[BG_NSLOT][BG_NW];
};

static_assert(sizeof(BgSmem) <= 160 * 1024 / 3, "three workgroups per CU");
static_assert(offsetof(BgSmem, cnt) % 4 == 0 && BG_W % 2 == 0, "cnt in whole 32-bit words");

// Adds one to counter p of cnt and returns its old value.  LDS atomics are
// 32-bit: the add goes to the word holding counters p & ~1 and p | 1; a
// counter never exceeds BG_W < 2^16, so no carry reaches its neighbour.
__device__ __forceinline__ int cnt_add(BgSmem& g, int p) {
    const int sh = (p & 1) * 16;
    return (int)((atomicAdd(reinterpret_cast<unsigned*>(g.cnt) + (p >> 1), 1u << sh) >> sh) & 0xFFFFu);
}

struct WsCols {
    int* col0;
    int* col1;
    double* val0;
    double* val1;
};

// Minimum over the wavefront (DPP prefix minimum, then the last lane).
__device__ __forceinline__ int wave_min(int x) {
    x = min(x, dpp_mov<0x111, 0xF>(x, INT_MAX));  // row_shr:1
    x = min(x, dpp_mov<0x112, 0xF>(x, INT_MAX));  // row_shr:2
    x = min(x, dpp_mov<0x114, 0xF>(x, INT_MAX));  // row_shr:4
    x = min(x, dpp_mov<0x118, 0xF>(x, INT_MAX));  // row_shr:8
    x = min(x, dpp_mov<0x142, 0xA>(x, INT_MAX));  // row_bcast15 into rows 1 and 3
    x = min(x, dpp_mov<0x143, 0xC>(x, INT_MAX));  // row_bcast31 into rows 2 and 3
    return wave_last(x);
}

// Block reductions and scans with one barrier each.  Every call takes the
// next of BG_NSLOT scratch slots (the slot counter is uniform), so a slot is
// written again only after several more barriers.
struct Slots {
    int i = 0;
    __device__ int next() {
        const int r = i;
        i = (i + 1) % BG_NSLOT;
        return r;
    }
};

__device__ __forceinline__ int block_sum_i(BgSmem& g, Slots& sl, int x) {
    const int k = sl.next();
    const int s = wave_sum(x);
    if (lane_id() == 0) g.red[k][threadIdx.x / BRM_WAVE] = s;
    __syncthreads();
    int tot = 0;
#pragma unroll
    for (int i = 0; i < BG_NW; ++i) tot += g.red[k][i];
    return tot;
}

__device__ __forceinline__ int64_t block_sum_l(BgSmem& g, Slots& sl, int64_t x) {
    const int k = sl.next();
    const int64_t s = wave_sum(x);
    if (lane_id() == 0) g.red64[k][threadIdx.x / BRM_WAVE] = s;
    __syncthreads();
    int64_t tot = 0;
#pragma unroll
    for (int i = 0; i < BG_NW; ++i) tot += g.red64[k][i];
    return tot;
}

__device__ __forceinline__ int block_min_i(BgSmem& g, Slots& sl, int x) {
    const int k = sl.next();
    const int m = wave_min(x);
    if (lane_id() == 0) g.red[k][threadIdx.x / BRM_WAVE] = m;
    __syncthreads();
    int r = INT_MAX;
#pragma unroll
    for (int i = 0; i < BG_NW; ++i) r = min(r, g.red[k][i]);
    return r;
}

__device__ __forceinline__ int block_scan_i(BgSmem& g, Slots& sl, int x, int* total) {
    return block_excl_scan_1b<BG_NT>(x, g.red[sl.next()], total);
}

// Buffer element i (0 <= i < BG_IPT) of the calling thread: wavefront w
// owns elements [w * IPT * 64, (w + 1) * IPT * 64), and element i of every
// lane of the wavefront is one contiguous run of 64 (conflict-free LDS
// accesses, coalesced loads).
__device__ __forceinline__ int bg_elem(int i) {
    return (int)(threadIdx.x / BRM_WAVE) * (BG_IPT * BRM_WAVE) + i * BRM_WAVE + lane_id();
}

// Children with at most BG_WHOLE elements left load all of them: a partly
// loaded child bounds the round at its last loaded column, and a short child
// given a proportional share of one or two elements would bound it early.
constexpr int BG_WHOLE = 16;
static_assert(BG_KMAX * BG_WHOLE <= BG_W - BG_KMAX, "whole short children fit the window");

// Window of child t (rem elements left, R left over all children) for the
// next round: everything left if it all fits; a short child's whole rest;
// else a share of what the short children leave of BG_W - BG_KMAX,
// proportional to what the child has left (at least one element).  Called by
// all of wavefront 0, which holds every child (k <= 64).
__device__ __forceinline__ int bg_share(int64_t rem, int64_t R) {
    const bool whole = rem > 0 && rem <= BG_WHOLE;
    const int64_t ws = wave_sum(whole ? rem : (int64_t)0);
    if (rem <= 0) return 0;
    if (R <= BG_W || whole) return (int)rem;
    int64_t x = (int64_t)((double)(BG_W - BG_KMAX - ws) * (double)rem / (double)(R - ws));
    if (x < 1) x = 1;
    if (x > rem) x = rem;
    return (int)x;
}

// Windows for the first round (bg_share).  Returns true when nothing is
// left.  The windows are published by the caller's next barrier.
__device__ __forceinline__ bool bg_windows(BgSmem& g, Slots& sl, int k) {
    static_assert(BG_KMAX <= BRM_WAVE, "every child in wavefront 0");
    const int t = threadIdx.x;
    const int64_t rem = t < k ? g.len[t] - g.cur[t] : 0;
    const int64_t R = block_sum_l(g, sl, rem);
    int wi = 0;
    if (t < BRM_WAVE) wi = bg_share(rem, R);
    // one scan gives the offsets (low 16 bits) and the rank among children
    // with a window (high bits)
    int tot;
    const int ex = block_scan_i(g, sl, wi | ((wi > 0 ? 1 : 0) << 16), &tot);
    if (t < k) {
        g.w[t] = wi;
        g.woff[t] = ex & 0xFFFF;
        g.fbase[t] = g.off[t] + g.cur[t] - (ex & 0xFFFF);
        if (wi > 0) {
            g.wchild[ex >> 16] = (unsigned char)t;
            atomicOr(&g.wstart[(ex & 0xFFFF) >> 6], 1ull << (ex & 63));
        }
    }
    if (t == BG_NT - 1) g.woff[k] = tot & 0xFFFF;
    return R == 0;
}

// Global source of a child's elements: a workspace buffer (a node of the
// previous pass) or a row of B read in place and scaled.
constexpr unsigned char SRC_WS0 = 0, SRC_WS1 = 1, SRC_B = 2;
struct SrcPtr {
    const int* c;
    const double* v;
    bool virt;
};
__device__ __forceinline__ SrcPtr src_ptr(unsigned char sk, const RowArgs& a, const WsCols& ws) {
    const bool virt = sk == SRC_B;
    return SrcPtr{virt ? a.bcol : (sk == SRC_WS1 ? ws.col1 : ws.col0),
                  virt ? a.bval : (sk == SRC_WS1 ? ws.val1 : ws.val0), virt};
}

// Loads this thread's elements of the proportional windows (buffer
// positions bg_elem(i)) into registers; rj = the element's child.
// The child of element e is the last window starting at or before e: its
// rank is the number of wstart bits up to e, less one (the 64 elements of
// one step of a wavefront share one wstart word).  Elements at or past the
// window end (Wn >= 1) load element Wn - 1 again and are not stored, so
// that nothing here branches and every batch of LDS reads is issued
// together.
__device__ __forceinline__ void bg_fetch(const BgSmem& g, int Wn, const RowArgs& a, const WsCols& ws,
                                         int (&rc)[BG_IPT], double (&rv)[BG_IPT], int (&rj)[BG_IPT],
                                         unsigned& vmask) {
    static_assert(BG_W / 64 <= BRM_WAVE, "one wstart word per lane");
    vmask = 0u;
    const int lane = lane_id();
    const int pc = lane < BG_W / 64 ? __popcll(g.wstart[lane]) : 0;
    const int before = wave_incl_scan(pc) - pc;  // window starts in earlier words
    const unsigned long long upto = lane == BRM_WAVE - 1 ? ~0ull : (2ull << lane) - 1ull;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / BRM_WAVE));
    int j[BG_IPT];
#pragma unroll
    for (int i = 0; i < BG_IPT; ++i) {
        const int wd = wv * BG_IPT + i;  // bg_elem(i) >> 6, the same for the whole wavefront
        j[i] = g.wchild[__builtin_amdgcn_readlane(before, wd) + __popcll(g.wstart[wd] & upto) - 1];
    }
#pragma unroll
    for (int i = 0; i < BG_IPT; ++i) {
        const SrcPtr p = src_ptr(g.src[j[i]], a, ws);
        const int64_t x = g.fbase[j[i]] + min(bg_elem(i), Wn - 1);
        rc[i] = p.c[x];
        rv[i] = p.v[x];
        rj[i] = j[i];
        if (p.virt) vmask |= 1u << i;
    }
}

// Sum of output bucket [s, s + m), 2 <= m <= BG_THREAD_M, by one thread in
// registers: the contributions come from distinct children, so sorting
// them by list id orders them as the tree's leaves; the sorted values are
// then folded in one pass: when the next leaf l follows leaf p and they
// first differ at bit d, every pending left subtree below height d on p's
// path is complete and is added (left + right) to p's side, which becomes
// the pending left sibling at height d; at the end all pending left subtrees
// are added in increasing height.  This is the reference's bracketing:
// siblings added left + right, a node with one present child passes it
// through.
__device__ double bg_small_sum(const BgSmem& g, int s, int m) {
    constexpr int M = BG_THREAD_M;
    // unconditional reads (indices clamped into the bucket), so that each
    // batch of LDS reads is issued together; slots past m get the top byte
    // set and sort last
    int ce[M];
#pragma unroll
    for (int i = 0; i < M; ++i) ce[i] = g.contrib[s + min(i, m - 1)];
    unsigned key[M];
#pragma unroll
    for (int i = 0; i < M; ++i)
        key[i] = ((unsigned)g.elid[ce[i]] << 16) | (unsigned)ce[i] | (i < m ? 0u : 0xFF000000u);
    // Batcher's odd-even merge sort network over M keys
#pragma unroll
    for (int p = 1; p < M; p <<= 1)
#pragma unroll
        for (int k = p; k >= 1; k >>= 1)
#pragma unroll
            for (int j = k % p; j + k < M; j += 2 * k)
#pragma unroll
                for (int i = 0; i < k; ++i)
                    if (i + j + k < M && (i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                        const unsigned x = key[i + j], y = key[i + j + k];
                        key[i + j] = x < y ? x : y;
                        key[i + j + k] = x < y ? y : x;
                    }
    double v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) v[i] = g.val[key[i] & 0xFFFFu];  // used for i < m only
    double lv[BPASS_BITS];
    unsigned pend = 0u;
    int prev = (int)(key[0] >> 16);
    double cur = v[0];
#pragma unroll
    for (int i = 1; i < M; ++i) {
        if (i < m) {
            const int l = (int)(key[i] >> 16);
            const int d = 31 - __clz(prev ^ l);
            double x = cur;
#pragma unroll
            for (int h = 0; h < BPASS_BITS; ++h)
                if (h < d && ((pend >> h) & 1u)) x = fadd(lv[h], x);
#pragma unroll
            for (int h = 0; h < BPASS_BITS; ++h)
                if (h == d) lv[h] = x;
            pend = (pend & ~((1u << d) - 1u)) | (1u << d);
            cur = v[i];
            prev = l;
        }
    }
    double x = cur;
#pragma unroll
    for (int h = 0; h < BPASS_BITS; ++h)
        if ((pend >> h) & 1u) x = fadd(lv[h], x);
    return x;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// a (left) and b (right) of one tree node: both present are added, one
// present passes through
__device__ __forceinline__ void node_combine(double& a, bool& pa, double b, bool pb) {
    if (pa && pb) a = fadd(a, b);
    else if (pb) a = b;
    pa = pa || pb;
}

// Value of lane (lane - shift) for the DPP control CTRL (row_shr:s within a
// row of 16 lanes, row_bcast15 / row_bcast31 across rows); lanes without a
// source keep 0.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const int lo = dpp_i<CTRL>(__double2loint(x));
    const int hi = dpp_i<CTRL>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// One tree level over lanes: lane i with i % (2 s) == 2 s - 1 holds the
// right node and takes the left one from lane i - s (CTRL moves it there).
template <int CTRL, int S>
__device__ __forceinline__ void dpp_level(double& x, bool& px, int lane) {
    double l = dpp_d<CTRL>(x);
    bool pl = dpp_i<CTRL>((int)px) != 0;
    if ((lane & (2 * S - 1)) == 2 * S - 1) {
        node_combine(l, pl, x, px);
        x = l;
        px = pl;
    }
}

// Sums of two output buckets by one wavefront, one per half of 32 lanes
// (m = 0: no bucket).  Each half places its bucket's contributions densely
// by list id (BG_KMAX slots of its own in dense/pres, pres all zero on entry
// and on return), each lane reduces its BG_KMAX / 32 slots, and the aligned tree over the
// half's lanes is finished with DPP moves (no LDS round trips): lanes
// 2s-1 mod 2s combine at level s, the roots end in lanes 31 and 63.
// Returns the first sum in r0 and the second in r1, in every lane.
__device__ void bg_wave_sum2(const BgSmem& g, int s0, int m0, int s1, int m1, double* dense,
                             unsigned char* pres, double& r0, double& r1) {
    const int lane = lane_id();
    const int h = lane >> 5, hl = lane & 31;
    const int s = h ? s1 : s0, m = h ? m1 : m0;
    double* dn = dense + h * BG_KMAX;
    unsigned char* pr = pres + h * BG_KMAX;
    for (int i = hl; i < m; i += 32) {
        const int e = g.contrib[s + i];
        const int l = g.elid[e];
        dn[l] = g.val[e];
        pr[l] = 1;
    }
    wave_lds_sync();
    constexpr int SPL = BG_KMAX / 32;  // slots per lane
    double v[SPL];
    bool p[SPL];
#pragma unroll
    for (int i = 0; i < SPL; ++i) {
        p[i] = pr[SPL * hl + i] != 0;
        v[i] = dn[SPL * hl + i];
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < SPL; ++i) pr[SPL * hl + i] = 0;
#pragma unroll
    for (int w = 1; w < SPL; w <<= 1)
#pragma unroll
        for (int i = 0; i + w < SPL; i += 2 * w) node_combine(v[i], p[i], v[i + w], p[i + w]);
    double x = v[0];
    bool px = p[0];
    dpp_level<0x111, 1>(x, px, lane);   // row_shr:1
    dpp_level<0x112, 2>(x, px, lane);   // row_shr:2
    dpp_level<0x114, 4>(x, px, lane);   // row_shr:4
    dpp_level<0x118, 8>(x, px, lane);   // row_shr:8
    dpp_level<0x142, 16>(x, px, lane);  // row_bcast15: lanes 31, 63 get 15, 47
    wave_lds_sync();
    r0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 31),
                          __builtin_amdgcn_readlane(__double2loint(x), 31));
    r1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 63),
                          __builtin_amdgcn_readlane(__double2loint(x), 63));
}

// At most 168 VGPRs (three waves per SIMD, launch bounds) and a third of the
// CU's LDS: three workgroups per CU, so that one workgroup's barrier and
// memory waits are covered by the other two.
__global__ __launch_bounds__(BG_NT, 3) void k_bgroup(const GroupTask* __restrict__ tasks,
                                                  const int* __restrict__ order,
                                                  const Child* __restrict__ children, RowArgs a, WsCols ws,
                                                  int64_t* __restrict__ slots, int64_t* counters, OutArgs o) {
    extern __shared__ __align__(16) unsigned char smraw[];
    BgSmem& g = *reinterpret_cast<BgSmem*>(smraw);
    const int t = threadIdx.x;
    const GroupTask T = tasks[order ? order[blockIdx.x] : (int)blockIdx.x];
    const int k = T.nchild;
    Slots sl;
    if (t < k) {
        const Child c = children[T.cbase + t];
        g.src[t] = c.src.kind == SRC_VIRTUAL ? SRC_B : c.src.buf ? SRC_WS1 : SRC_WS0;
        g.lid[t] = (unsigned char)c.lid;
        g.off[t] = c.src.off;
        g.scale[t] = c.src.scale;
        g.len[t] = c.src.kind == SRC_VIRTUAL ? c.src.len : slots[c.src.slot];
        g.cur[t] = 0;
    }
    if (t < 32)
        for (int it = 0; it < BG_TOPUPS; ++it) g.hist[it][t] = 0;
    if (t < BG_W / 64) {
        g.wstart[t] = 0ull;
        g.xstart[t] = 0ull;
    }
    if (t == 0) {
        g.nbig = 0;
        g.nmid = 0;
        g.bnext = 0;
    }
#pragma unroll
    for (int i = 0; i < BG_WPT; ++i) g.bm[t * BG_WPT + i] = 0ull;
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
    const int64_t in_total = block_sum_l(g, sl, t < k ? g.len[t] : 0);

    int rc[BG_IPT], rj[BG_IPT];
    double rv[BG_IPT];
    unsigned vmask = 0u;
    bool done = bg_windows(g, sl, k);
    __syncthreads();
    int Wn = g.woff[k];
    if (!done) bg_fetch(g, Wn, a, ws, rc, rv, rj, vmask);
    int64_t out_n = 0;
    int rounds = 0;
    // shader cycles per round phase, seen by thread 0 (barrier waits included)
    long long tph[N_PHASES] = {};
    long long tprev = clock64();
    auto mark = [&](int p) {
        if (t == 0) {
            const long long now = clock64();
            tph[p] += now - tprev;
            tprev = now;
        }
    };
    int64_t R = in_total;  // elements not yet taken
    while (!done) {
        rounds++;
        // every load of the prefetched windows is complete here (vmcnt 0):
        // stated explicitly so that no load is left in flight into the next
        // prefetch, whose destination registers are reused as temporaries
        __builtin_amdgcn_s_waitcnt(0x0F70);
        // ---- the proportional windows into the buffer (products rounded
        // here; the children's scales and list ids read for every element,
        // so that the reads are issued together)
        {
            double sc[BG_IPT];
            int lq[BG_IPT];
#pragma unroll
            for (int i = 0; i < BG_IPT; ++i) {
                sc[i] = g.scale[rj[i]];
                lq[i] = g.lid[rj[i]];
            }
#pragma unroll
            for (int i = 0; i < BG_IPT; ++i) {
                const int e = bg_elem(i);
                if (e < Wn) {
                    g.col[e] = rc[i];
                    g.val[e] = ((vmask >> i) & 1u) ? fmul(sc[i], rv[i]) : rv[i];
                    g.elid[e] = (unsigned char)lq[i];
                }
            }
        }
        __syncthreads();
        // last loaded column of every child with a window (read by the
        // child's own thread until the next barrier)
        if (t < k && g.w[t] > 0) g.lastc[t] = g.col[g.woff[t] + g.w[t] - 1];
        // the window starts were read by the fetch of this round's windows,
        // before the barrier above; the next windows set them in phase 5
        if (t < BG_W / 64) g.wstart[t] = 0ull;
        const int64_t rem = t < k ? g.len[t] - g.cur[t] : 0;
        const int c_lo = block_min_i(g, sl, (t < k && g.w[t] > 0) ? g.col[g.woff[t]] : INT_MAX);
        int filled = Wn;
        // this child's loaded elements: its window, then one run per top-up
        // (buffer offset and length), columns ascending across them
        int sgo[1 + BG_TOPUPS], sgn[1 + BG_TOPUPS];
        sgo[0] = t < k ? g.woff[t] : 0;
        sgn[0] = t < k ? g.w[t] : 0;
#pragma unroll
        for (int it = 0; it < BG_TOPUPS; ++it) sgo[1 + it] = sgn[1 + it] = 0;
        mark(0);
        // ---- top-ups: children stopping short of the target share the free space
#pragma unroll
        for (int it = 0; it < BG_TOPUPS; ++it) {
            const bool partial = t < k && g.w[t] > 0 && (int64_t)g.w[t] < rem;
            // log2 of the extent lastc - c_lo + 1 (>= 1) loaded so far
            const int bkt = partial ? 31 - __clz(g.lastc[t] - c_lo + 1) : 0;
            if (partial) atomicAdd(&g.hist[it][bkt], 1);
            __syncthreads();
            // from one wave scan of the histogram: the number of partial
            // children, the median bucket mb, and the number of children
            // below it (extent < 2^mb: short of the target c_lo + 2^mb - 1)
            int np, mb, ns;
            {
                const int lane = lane_id();
                const int hv = lane < 32 ? g.hist[it][lane] : 0;
                const int inc = wave_incl_scan(hv);
                np = __builtin_amdgcn_readlane(inc, 31);
                const uint64_t m = __ballot(inc >= (np + 1) / 2);
                mb = np > 0 ? __ffsll((long long)m) - 1 : 0;
                ns = __builtin_amdgcn_readlane(inc - hv, mb);
            }
            if (np == 0 || filled >= BG_W) break;
            const bool shorty = partial && bkt < mb;
            const int add = ns > 0 ? (BG_W - filled) / ns : 0;
            if (add == 0) break;
            const int amt = shorty ? (int)min((int64_t)add, rem - g.w[t]) : 0;
            // one scan gives the run offsets (low 16 bits) and the rank
            // among children with a run (high bits)
            int tot;
            const int ex = block_scan_i(g, sl, amt | ((amt > 0 ? 1 : 0) << 16), &tot);
            const int xo = ex & 0xFFFF;
            tot &= 0xFFFF;
            if (amt > 0) {
                g.xchild[ex >> 16] = (unsigned char)t;
                g.xbase[t] = g.off[t] + g.cur[t] + g.w[t] - xo;
                atomicOr(&g.xstart[xo >> 6], 1ull << (xo & 63));
            }
            __syncthreads();
            // tot <= BG_W - filled: at most BG_IPT top-up elements per
            // thread (element t + i * BG_NT; the 64 of one wavefront step
            // share one xstart word).  As in the window fetch, the child of
            // an element is found by a popcount over the run starts, and
            // elements at or past tot (>= 1) load element tot - 1 again and
            // are not stored, so that the LDS reads are issued together;
            // wavefront steps entirely past tot load nothing.
            int tc[BG_IPT], tl[BG_IPT];
            double tv[BG_IPT], ts[BG_IPT];
            unsigned tvm = 0u;
            {
                const int lane = lane_id();
                const int pc = lane < BG_W / 64 ? __popcll(g.xstart[lane]) : 0;
                const int before = wave_incl_scan(pc) - pc;
                const unsigned long long upto = lane == BRM_WAVE - 1 ? ~0ull : (2ull << lane) - 1ull;
                const int wv = __builtin_amdgcn_readfirstlane((int)(t / BRM_WAVE));
                int j[BG_IPT];
#pragma unroll
                for (int i = 0; i < BG_IPT; ++i) {
                    const int wd = wv + i * BG_NW;  // (t + i * BG_NT) >> 6
                    j[i] = g.xchild[__builtin_amdgcn_readlane(before, wd) + __popcll(g.xstart[wd] & upto) - 1];
                }
#pragma unroll
                for (int i = 0; i < BG_IPT; ++i) {
                    const SrcPtr p = src_ptr(g.src[j[i]], a, ws);
                    const int64_t x = g.xbase[j[i]] + min(t + i * BG_NT, tot - 1);
                    ts[i] = g.scale[j[i]];
                    tl[i] = g.lid[j[i]];
                    if (p.virt) tvm |= 1u << i;
                    tc[i] = 0;
                    tv[i] = 0.0;
                    if (wv * BRM_WAVE + i * BG_NT < tot) {  // the same for the whole wavefront
                        tc[i] = p.c[x];
                        tv[i] = p.v[x];
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < BG_IPT; ++i) {
                const int e = t + i * BG_NT;
                if (e < tot) {
                    const int pos = filled + e;
                    g.col[pos] = tc[i];
                    g.val[pos] = ((tvm >> i) & 1u) ? fmul(ts[i], tv[i]) : tv[i];
                    g.elid[pos] = (unsigned char)tl[i];
                }
            }
            __syncthreads();
            if (t < BG_W / 64) g.xstart[t] = 0ull;  // read by every thread before the barrier above
            if (amt > 0) {
                g.lastc[t] = g.col[filled + xo + amt - 1];
                g.w[t] += amt;
                sgo[1 + it] = filled + xo;
                sgn[1 + it] = amt;
            }
            filled += tot;
        }
        mark(1);
        // ---- round bound
        const bool partial = t < k && g.w[t] > 0 && (int64_t)g.w[t] < rem;
        const int lm = block_min_i(g, sl, partial ? g.lastc[t] : INT_MAX);
        const int bound = (int)min((int64_t)lm, (int64_t)c_lo + BG_SPAN - 1);
        // every read of the histograms was before the barrier above
        if (t < 32)
#pragma unroll
            for (int it = 0; it < BG_TOPUPS; ++it) g.hist[it][t] = 0;
        // ---- taken elements: bitmap bits (kept: 0 for the others); counters cleared
        unsigned taken = 0u;
        int bits[BG_IPT];
#pragma unroll
        for (int i = 0; i < BG_IPT; ++i) {
            const int e = bg_elem(i);
            const int c = g.col[e];
            bits[i] = 0;
            if (e < filled && c <= bound) {
                taken |= 1u << i;
                bits[i] = c - c_lo;
                atomicOr(&g.bm[bits[i] >> 6], 1ull << (bits[i] & 63));
            }
            g.cnt[e] = 0;
        }
        // ---- this child's taken count: its elements with column <= bound,
        // a prefix of its runs (binary search in the first run that ends
        // above the bound)
        int64_t remt = 0;
        if (t < k) {
            int ntk = 0;
#pragma unroll
            for (int r = 0; r <= BG_TOPUPS; ++r) {
                const int o = sgo[r], n = sgn[r];
                if (n == 0) continue;
                if (g.col[o + n - 1] <= bound) {
                    ntk += n;
                    continue;
                }
                // the first element above the bound is in [lo, hi]; eight
                // probes read together narrow it to about a ninth first
                int lo = 0, hi = n - 1;  // col[o + hi] > bound
                if (n > 16) {
                    int pv[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) pv[j] = g.col[o + (j + 1) * (n - 1) / 9];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int pj = (j + 1) * (n - 1) / 9;
                        if (pv[j] <= bound) lo = pj + 1;
                    }
#pragma unroll
                    for (int j = 7; j >= 0; --j) {
                        const int pj = (j + 1) * (n - 1) / 9;
                        if (pv[j] > bound) hi = pj;
                    }
                }
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (g.col[o + mid] <= bound) lo = mid + 1;
                    else hi = mid;
                }
                ntk += lo;
                break;
            }
            g.cur[t] += ntk;
            remt = g.len[t] - g.cur[t];
        }
        __syncthreads();
        mark(2);
        // ---- ranks of the bitmap words; with them the number taken
        int pc[BG_WPT];
        int sum = 0;
#pragma unroll
        for (int i = 0; i < BG_WPT; ++i) {
            pc[i] = __popcll(g.bm[t * BG_WPT + i]);
            sum += pc[i];
        }
        int nout;
        {
            int tot;
            const int ex = block_scan_i(g, sl, sum | (__popc(taken) << 16), &tot);
            nout = tot & 0xFFFF;
            R -= tot >> 16;
            int wex = ex & 0xFFFF;
#pragma unroll
            for (int i = 0; i < BG_WPT; ++i) {
                g.wrank[t * BG_WPT + i] = (unsigned short)wex;
                wex += pc[i];
            }
        }
        done = R == 0;
        __syncthreads();
        mark(3);
        // ---- output position of every taken element; contributions per position
        // (read for every element, so that the reads are issued together)
        int pos[BG_IPT];
#pragma unroll
        for (int i = 0; i < BG_IPT; ++i) {
            const int wd = bits[i] >> 6;
            pos[i] = g.wrank[wd] + __popcll(g.bm[wd] & ((1ull << (bits[i] & 63)) - 1ull));
        }
#pragma unroll
        for (int i = 0; i < BG_IPT; ++i)
            if ((taken >> i) & 1u) (void)cnt_add(g, pos[i]);
        __syncthreads();
        mark(4);
        // ---- bucket offsets (exclusive scan of the counts) and, in the
        // same scan, the next round's windows (bg_share)
        {
            int wi = 0;
            if (t < BRM_WAVE) wi = bg_share(remt, R);
            int loc[BG_IPT];
            int sloc = 0;
#pragma unroll
            for (int i = 0; i < BG_IPT; ++i) {
                const int oi = t * BG_IPT + i;
                loc[i] = oi < nout ? g.cnt[oi] : 0;
                sloc += loc[i];
            }
            // bits 0-15: bucket offset, 16-31: window offset, 32-: window rank
            const int64_t packed = (int64_t)sloc | ((int64_t)wi << 16) | ((int64_t)(wi > 0 ? 1 : 0) << 32);
            int64_t tot;
            const int64_t ex = block_excl_scan_1b<BG_NT>(packed, g.red64[sl.next()], &tot);
            int bex = (int)(ex & 0xFFFF);
#pragma unroll
            for (int i = 0; i < BG_IPT; ++i) {
                const int oi = t * BG_IPT + i;
                if (oi < nout) g.cnt[oi] = (unsigned short)bex;
                bex += loc[i];
            }
            const int wo = (int)((ex >> 16) & 0xFFFF);
            if (t < k) {
                g.w[t] = wi;
                g.woff[t] = wo;
                g.fbase[t] = g.off[t] + g.cur[t] - wo;
                if (wi > 0) {
                    g.wchild[(int)(ex >> 32)] = (unsigned char)t;
                    atomicOr(&g.wstart[wo >> 6], 1ull << (wo & 63));
                }
            }
            Wn = (int)((tot >> 16) & 0xFFFF);
            if (t == BG_NT - 1) g.woff[k] = Wn;
        }
        __syncthreads();
        mark(5);
        // ---- bucket the contributions (cnt[p] ends as the bucket's end);
        // the next round's windows are fetched into registers
        int slot[BG_IPT];
#pragma unroll
        for (int i = 0; i < BG_IPT; ++i) {
            slot[i] = 0;
            if ((taken >> i) & 1u) slot[i] = cnt_add(g, pos[i]);
        }
#pragma unroll
        for (int i = 0; i < BG_IPT; ++i)
            if ((taken >> i) & 1u) g.contrib[slot[i]] = (unsigned short)bg_elem(i);
        if (!done) bg_fetch(g, Wn, a, ws, rc, rv, rj, vmask);
        __syncthreads();
        mark(6);
        // ---- one output per thread (strided: coalesced stores) for one
        // or two contributions; outputs with more are queued, so that the
        // lanes of a wavefront take the same path: up to BG_THREAD_M
        // contributions for a second thread pass, more for whole wavefronts
        for (int oi = t; oi < nout; oi += BG_NT) {
            const int sp = g.cnt[oi > 0 ? oi - 1 : 0];
            const int s = oi > 0 ? sp : 0;
            const int m = g.cnt[oi] - s;
            if (m > 2) {
                if (m > BG_THREAD_M) g.big[atomicAdd(&g.nbig, 1)] = (unsigned short)oi;
                else g.mid[atomicAdd(&g.nmid, 1)] = (unsigned short)oi;
                continue;
            }
            // both reads in every lane (e1 == e0 for one contribution), so
            // that the LDS reads are issued together
            const int e0 = g.contrib[s];
            const int e1 = g.contrib[s + m - 1];
            const double v0 = g.val[e0], v1 = g.val[e1];
            const unsigned l0 = g.elid[e0], l1 = g.elid[e1];
            double v = v0;
            if (m == 2) v = l0 < l1 ? fadd(v0, v1) : fadd(v1, v0);
            const int64_t q = out_n + oi;
            if (q < dcap) {
                dc[q] = g.col[e0];
                dv[q] = v;
            }
        }
        __syncthreads();
        {
            const int nmid = g.nmid;
            for (int b = t; b < nmid; b += BG_NT) {
                const int oi = g.mid[b];
                const int s = oi ? g.cnt[oi - 1] : 0;
                const double v = bg_small_sum(g, s, g.cnt[oi] - s);
                const int64_t q = out_n + oi;
                if (q < dcap) {
                    dc[q] = g.col[g.contrib[s]];
                    dv[q] = v;
                }
            }
        }
        {
            // the bitmap and the word ranks are free until the next round:
            // wave w sums two buckets at a time densely in the first 512
            // words of its share of the bitmap, with flags in bytes
            // [512 w, 512 w + 512) of the ranks, then clears its share of the
            // bitmap.  The pairs are claimed from a shared counter, so the
            // waves with no second-pass buckets above take more of them.
            static_assert(2 * BG_KMAX <= BG_WORDS / BG_NW, "dense sums fit a wave's bitmap share");
            static_assert(BG_NW * 2 * BG_KMAX <= (int)sizeof(BgSmem::wrank), "flags fit the word ranks");
            const int wv = t / BRM_WAVE, lane = lane_id();
            double* dense = reinterpret_cast<double*>(&g.bm[wv * (BG_WORDS / BG_NW)]);
            unsigned char* pres = reinterpret_cast<unsigned char*>(g.wrank) + wv * 2 * BG_KMAX;
            const int nbig = g.nbig;
            if (nbig > 0) {
                static_assert((2 * BG_KMAX) % 4 == 0, "whole flag words");
#pragma unroll
                for (int i = lane; i < 2 * BG_KMAX / 4; i += BRM_WAVE) reinterpret_cast<unsigned*>(pres)[i] = 0u;
                wave_lds_sync();
                for (;;) {
                    int b = 0;
                    if (lane == 0) b = atomicAdd(&g.bnext, 2);
                    b = __builtin_amdgcn_readfirstlane(b);  // lane 0's claim
                    if (b >= nbig) break;
                    const int oi0 = g.big[b];
                    const int oi1 = b + 1 < nbig ? g.big[b + 1] : -1;
                    const int s0 = oi0 ? g.cnt[oi0 - 1] : 0;
                    const int s1 = oi1 > 0 ? g.cnt[oi1 - 1] : 0;
                    const int m1 = oi1 >= 0 ? g.cnt[oi1] - s1 : 0;
                    double v0, v1;
                    bg_wave_sum2(g, s0, g.cnt[oi0] - s0, s1, m1, dense, pres, v0, v1);
                    const int oi = lane == 0 ? oi0 : oi1;
                    const int64_t q = out_n + oi;
                    if ((lane == 0 || (lane == 32 && oi1 >= 0)) && q < dcap) {
                        dc[q] = g.col[g.contrib[lane == 0 ? s0 : s1]];
                        dv[q] = lane == 0 ? v0 : v1;
                    }
                }
            }
            static_assert(BG_WPT * BRM_WAVE == BG_WORDS / BG_NW, "a wave clears its own bitmap words");
#pragma unroll
            for (int i = 0; i < BG_WPT; ++i) g.bm[t * BG_WPT + i] = 0ull;
        }
        out_n += nout;
        __syncthreads();
        if (t == 0) {  // read by every wave before the barrier above
            g.nbig = 0;
            g.nmid = 0;
            g.bnext = 0;
        }
        mark(7);
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
#pragma unroll
        for (int p = 0; p < N_PHASES; ++p)
            atomicAdd((unsigned long long*)&counters[CNT_PH0 + p], (unsigned long long)tph[p]);
    }
}

}  // namespace

hipError_t launch_bgroup(const GroupTask* tasks, int64_t n, const Child* children, const RowArgs& a,
                         int* const ws_col[2], double* const ws_val[2], int64_t* slots, int64_t* counters,
                         const OutArgs& o, hipStream_t s, const int* order) {
    if (n <= 0) return hipSuccess;
    static SmemLimit lim;
    if (hipError_t e = lim.raise((const void*)k_bgroup, (int)sizeof(BgSmem)); e != hipSuccess) return e;
    WsCols w{ws_col[0], ws_col[1], ws_val[0], ws_val[1]};
    {
        ProfScope ps_("k_bgroup", s);
        const int64_t ch = launch_chunk(BG_NT);
        for (int64_t b = 0; b < n; b += ch)
            k_bgroup<<<(unsigned)std::min<int64_t>(ch, n - b), BG_NT, sizeof(BgSmem), s>>>(
                order ? tasks : tasks + b, order ? order + b : nullptr, children, a, w, slots, counters, o);
    }
    return hipGetLastError();
}

}  // namespace brm
