// Hash-SpGEMM numeric tier (the paper's comparison baseline,
// Existing-SpGEMM/inc/hash_mult.h:558-608, binning BIN.h:157-177), for rows
// of at most HASH_CAP products: a selectable comparator to the BRMerge
// tiers, enabled by the hash_cap knob.
//
// One wavefront per row, no barriers.  The row's columns go into an LDS
// open-addressing table (hash (key * 107) & (size - 1), linear probing, as
// hash_mult.h:586-601).  Products are visited in the reference's order --
// A entries in order, each B row in order -- 64 at a time, and each one is
// added onto its column's value as t_val + value (:588), so each value is
// the sequential left fold of its column's contributions in list order,
// bit for bit the reference's.  Inside one step of 64 products, lanes that
// hit the same column commit in lane order (an LDS atomicMin elects the
// lowest pending lane of each column per sub-step).  The occupied slots are
// then sorted by column (sortOutput, :519-552) with a wavefront bitonic sort
// in LDS and written out.
//
// Each row's table has the smallest power of two of slots that is at least
// twice its product count (at least 64).  Size and layout never change a
// value, so they need not follow the reference's MIN_HT_N << (bin - 1).
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr unsigned HT_NONE = 0xFFFFu;    // no owner / no lane
constexpr unsigned HASH_SCAL = 107u;     // hash_mult.h:23

// LDS of one row of at most CAP products and CAP lists.
template <int CAP>
struct HashSmem {
    static constexpr int SLOTS = 2 * CAP;  // largest table (power of two)
    double val[SLOTS];
    int key[SLOTS];
    unsigned own[SLOTS];  // bits 0-15: electing lane (HT_NONE: none); bit 16: has a value
    int loff[CAP + 1];    // product offset of each list in the row; reused as sort keys
    int bst[CAP];         // B row start of each list; reused as sort payloads (slots)
};

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int CAP>
__global__ __launch_bounds__(BRM_WAVE) void k_hash_row(const int* __restrict__ rows, int nrows, RowArgs a, OutArgs o) {
    extern __shared__ __align__(16) unsigned char smraw[];
    HashSmem<CAP>& g = *reinterpret_cast<HashSmem<CAP>*>(smraw);
    const int lane = lane_id();
    const int row = rows[blockIdx.x];
    const int ab = a.arpt[row];
    const int L = a.arpt[row + 1] - ab;
    const int64_t ib = a.ipA[ab];
    const int S = (int)(a.ipA[ab + L] - ib);  // products, <= CAP
    int hs = BRM_WAVE;
    while (hs < 2 * S) hs <<= 1;  // <= 2 * CAP
    for (int k = lane; k < hs; k += BRM_WAVE) {
        g.key[k] = -1;
        g.own[k] = HT_NONE;
    }
    for (int k = lane; k <= L; k += BRM_WAVE) {
        g.loff[k] = (int)(a.ipA[ab + k] - ib);
        if (k < L) g.bst[k] = a.brpt[a.acol[ab + k]];
    }
    wsync();
    int hint = 0;
    for (int e0 = 0; e0 < S; e0 += BRM_WAVE) {
        const int e = e0 + lane;
        bool pending = e < S;
        int col = 0, slot = 0;
        double v = 0.0;
        if (pending) {
            hint = gallop_idx(g.loff, hint, L, e);
            const int64_t k = (int64_t)g.bst[hint] + (e - g.loff[hint]);
            col = a.bcol[k];
            v = fmul(a.aval[ab + hint], a.bval[k]);
            unsigned h = ((unsigned)col * HASH_SCAL) & (unsigned)(hs - 1);
            for (;;) {
                const int old = atomicCAS(&g.key[h], -1, col);
                if (old == -1 || old == col) break;
                h = (h + 1) & (unsigned)(hs - 1);
            }
            slot = (int)h;
        }
        // commit in lane (= list) order among the lanes of one column
        while (__ballot(pending) != 0ull) {
            if (pending) atomicMin(&g.own[slot], (g.own[slot] & 0x10000u) | (unsigned)lane);
            wsync();
            if (pending && (g.own[slot] & 0xFFFFu) == (unsigned)lane) {
                const unsigned st = g.own[slot];
                g.val[slot] = (st & 0x10000u) ? fadd(v, g.val[slot]) : v;
                g.own[slot] = 0x10000u | HT_NONE;
                pending = false;
            }
            wsync();
        }
    }
    // occupied slots (key, slot), compacted in slot order
    int n = 0;
    for (int k0 = 0; k0 < hs; k0 += BRM_WAVE) {
        const int k = k0 + lane;
        const int kk = g.key[k];
        const uint64_t m = __ballot(kk != -1);
        if (kk != -1) {
            const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
            g.loff[pos] = kk;
            g.bst[pos] = k;
        }
        n += __popcll(m);
    }
    wsync();
    // bitonic sort of n (<= CAP) keys, padded to a power of two
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = n + lane; k < np2; k += BRM_WAVE) {
        g.loff[k] = INT_MAX;
        g.bst[k] = -1;
    }
    wsync();
    for (int size = 2; size <= np2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < np2 / 2; i += BRM_WAVE) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const int kl = g.loff[lo], kh = g.loff[hi];
                if ((kl > kh) == up) {
                    g.loff[lo] = kh;
                    g.loff[hi] = kl;
                    const int t = g.bst[lo];
                    g.bst[lo] = g.bst[hi];
                    g.bst[hi] = t;
                }
            }
            wsync();
        }
    }
    int* dcol;
    double* dval;
    if (!o.upper) {
        const int64_t cb = o.crpt[row];
        const int64_t cap = o.crpt[row + 1] - cb;
        if (lane == 0 && (int64_t)n != cap) atomicOr(o.err, (int)DERR_COUNT_MISMATCH);
        if ((int64_t)n > cap) n = (int)cap;
        dcol = o.ccol + cb;
        dval = o.cval + cb;
    } else {
        dcol = o.ccol + ib;
        dval = o.cval + ib;
        if (lane == 0) o.rowlen[row] = n;
    }
    for (int i = lane; i < n; i += BRM_WAVE) {
        dcol[i] = g.loff[i];
        dval[i] = g.val[g.bst[i]];
    }
}

}  // namespace

namespace {
template <int CAP>
hipError_t hash_numeric_t(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s) {
    static SmemLimit lim;
    if (hipError_t e = lim.raise((const void*)k_hash_row<CAP>, (int)sizeof(HashSmem<CAP>)); e != hipSuccess) return e;
    {
        ProfScope ps_("k_hash_row", s);
        const int64_t ch = launch_chunk(BRM_WAVE);
        for (int64_t b = 0; b < n; b += ch) {
            const int c = (int)std::min<int64_t>(ch, n - b);
            k_hash_row<CAP><<<c, BRM_WAVE, sizeof(HashSmem<CAP>), s>>>(rows + b, c, a, o);
        }
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_hash_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap) {
    if (n <= 0) return hipSuccess;
    // a small table keeps many rows resident per CU
    return cap <= HASH_SMALL ? hash_numeric_t<HASH_SMALL>(rows, n, a, o, s) : hash_numeric_t<HASH_CAP>(rows, n, a, o, s);
}

}  // namespace brm
