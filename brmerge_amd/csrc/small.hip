// Short-row products in two launches and one host round trip.
//
// A product whose rows are all short (at most G = 64, or 32, lists and
// products) spends more on the general pipeline's dozen launches and two
// host round trips than on its kernels.  Such a product runs as:
//   1. one fill clearing the state: flags, counters, workgroup tickets and
//      look-back states;
//   2. k_small<G>: each workgroup takes a ticket (tickets are handed out in
//      the order workgroups start) and with it a run of consecutive rows.
//      One group of G lanes per row checks the row's own indices, finds its
//      list offsets by a group scan (no global product scan) and merges the
//      row in registers with the short tier's code (short_merge.h).  The
//      workgroup's nnz is scanned across workgroups by decoupled look-back
//      (wavefront 0 reads 64 predecessors' states at a time; a workgroup
//      waits only on workgroups that started before it), and every row is
//      written straight from registers into C at its place, with C's row
//      pointer.  Workgroups past the row workgroups check all of B (row
//      pointers, column range, sorted rows), as the general validation does.
// C is allocated beforehand for G entries per row, a bound of nnz(C) when
// every row is short.  The host reads the state once, at the end: a row with
// more than G lists or products sets a flag, and the product then goes
// through the general pipeline instead.
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"
#include "short_merge.h"

namespace brm {
namespace {

constexpr int SM_NT = 1024;
constexpr int SM_WPB = SM_NT / BRM_WAVE;

// Look-back states: bits 62-63 the state (1: aggregate, 2: inclusive
// prefix), bits 0-61 the value.  Each state word carries its own value, so
// relaxed atomics (no cache invalidation or write-back) are enough.
constexpr uint64_t LB_AGG = 1ull << 62, LB_INC = 2ull << 62, LB_VAL = (1ull << 62) - 1ull;
// States read per lane and look-back step.  The inclusive prefixes move
// forward by at most one window (64 * SM_LBW workgroups) per memory round
// trip, which bounds the kernel's rate in workgroups per microsecond.
constexpr int SM_LBW = 1;

// Exclusive prefix of this workgroup's `agg` over the workgroups before
// ticket b; called by all of wavefront 0.  Lane l reads the states at
// distances l * SM_LBW + j + 1 (j < SM_LBW) below the window's top, so that
// position l * SM_LBW + j orders them from nearest to farthest.
__device__ int64_t look_back(uint64_t* lb, int b, int64_t agg) {
    const int lane = lane_id();
    int64_t prefix = 0;
    if (b > 0) {
        if (lane == 0) __hip_atomic_store(&lb[b], LB_AGG | (uint64_t)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int hi = b - 1;;) {
            uint64_t v[SM_LBW];
#pragma unroll
            for (int j = 0; j < SM_LBW; ++j) {
                const int q = hi - lane * SM_LBW - j;
                v[j] = q >= 0 ? __hip_atomic_load(&lb[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : LB_INC;  // before workgroup 0: an empty prefix
            }
            int fj = SM_LBW;  // this lane's nearest inclusive prefix
#pragma unroll
            for (int j = SM_LBW - 1; j >= 0; --j)
                if ((v[j] & ~LB_VAL) == LB_INC) fj = j;
            const uint64_t incm = __ballot(fj < SM_LBW);
            const int first = incm ? __ffsll((long long)incm) - 1 : BRM_WAVE;  // lane of the nearest one
            const int ffj = __shfl(fj, first & (BRM_WAVE - 1), BRM_WAVE);
            const int lim = lane < first ? SM_LBW : (lane == first ? ffj + 1 : 0);
            bool unpublished = false;
            int64_t sum = 0;
#pragma unroll
            for (int j = 0; j < SM_LBW; ++j) {
                if (j < lim) {
                    unpublished |= v[j] == 0;
                    sum += (int64_t)(v[j] & LB_VAL);
                }
            }
            if (__ballot(unpublished) != 0ull) {  // a state not yet published
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += wave_sum(sum);
            if (first < BRM_WAVE) break;
            hi -= BRM_WAVE * SM_LBW;
        }
    }
    if (lane == 0)
        __hip_atomic_store(&lb[b], LB_INC | (uint64_t)(prefix + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

template <int G>
__global__ __launch_bounds__(SM_NT) void k_small(RowArgs a, int64_t nnzA, int64_t nnzB, int row_blocks,
                                                 int64_t* __restrict__ crpt, int* __restrict__ ccol,
                                                 double* __restrict__ cval, SmallState* __restrict__ st,
                                                 uint64_t* __restrict__ lb) {
    constexpr int RPW = BRM_WAVE / G;       // rows per wavefront
    constexpr int RPB = SM_WPB * RPW;       // rows per workgroup
    static_assert(RPB <= BRM_WAVE, "wavefront 0 scans the workgroup's rows");
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / BRM_WAVE);
    if ((int)blockIdx.x >= row_blocks) {
        // ---- B: one row per wavefront
        const int r = ((int)blockIdx.x - row_blocks) * SM_WPB + w;
        if (r > a.K) return;
        bool bad = false;
        if (r == a.K) {
            if (lane == 0) bad = a.brpt[0] != 0 || (int64_t)a.brpt[a.K] != nnzB;
        } else {
            const int b0 = a.brpt[r], b1 = a.brpt[r + 1];
            if (b0 < 0 || b0 > b1 || (int64_t)b1 > nnzB) {
                bad = true;
            } else {
                for (int k = b0 + lane; k < b1; k += BRM_WAVE) {
                    const int c = a.bcol[k];
                    bad |= c < 0 || c >= a.N || (k + 1 < b1 && c >= a.bcol[k + 1]);
                }
            }
        }
        if (__ballot(bad) != 0ull && lane == 0) atomicOr(&st->err, (int)DERR_INVALID);
        return;
    }
    __shared__ int ticket;
    __shared__ int rcnt[RPB];      // nnz of the workgroup's rows
    __shared__ int rip[RPB];       // their products
    __shared__ int64_t rbase[RPB];  // their C offsets
    if (threadIdx.x == 0) ticket = atomicAdd(&st->ticket, 1);
    __syncthreads();
    const int b = ticket;
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint64_t gmask = G == BRM_WAVE ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const int slot = w * RPW + lane / G;  // the row's place in the workgroup
    const int row = b * RPB + slot;
    const bool has = row < a.M;
    // ---- the row's own indices, checked before they are followed
    bool bad = false, miss = false;
    int a0 = 0, L = 0;
    if (has) {
        a0 = a.arpt[row];
        const int a1 = a.arpt[row + 1];
        if (a0 < 0 || a0 > a1 || (int64_t)a1 > nnzA) bad = true;
        else L = a1 - a0;
        if (row == 0 && a0 != 0) bad = true;
        if (row == a.M - 1 && (int64_t)a1 != nnzA) bad = true;
    }
    if (L > G) {
        miss = true;
        L = 0;
    }
    int lbst = 0;
    int64_t blen = 0;
    double lav = 0.0;
    if (!bad && gl < L) {
        const int c = a.acol[a0 + gl];
        if (c < 0 || c >= a.K) {
            bad = true;
        } else {
            const int b0 = a.brpt[c], b1 = a.brpt[c + 1];
            if (b0 < 0 || b0 > b1 || (int64_t)b1 > nnzB) {
                bad = true;
            } else {
                lbst = b0;
                blen = b1 - b0;
            }
        }
        lav = a.aval[a0 + gl];
    }
    bad = (__ballot(bad) & gmask) != 0ull;
    if (bad) blen = 0;
    // list offsets: inclusive scan of the list lengths within the group
    int64_t incl = wave_incl_scan(blen);
    if (G < BRM_WAVE) incl -= __shfl(incl, (gbase + BRM_WAVE - 1) & (BRM_WAVE - 1), BRM_WAVE) * (gbase > 0);
    const int64_t ip = __shfl(incl, gbase + G - 1, BRM_WAVE);
    if (ip > G) miss = true;
    const bool ok = has && !bad && !miss;
    const int off = (gl < L) ? (int)(incl - blen) : INT_MAX;
    const int ipn = ok ? (int)ip : 0;
    // product lane -> its list (largest p with off[p] <= gl), as the short tier
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
    const bool act = gl < ipn;
    int col = INT_MAX;
    double val = 0.0;
    if (act) {
        const int bi = p_bst + (gl - p_off);
        col = a.bcol[bi];
        val = fmul(p_av, a.bval[bi]);
    }
    // columns in range and strictly increasing within each list
    const int prev = __shfl(col, (lane + BRM_WAVE - 1) & (BRM_WAVE - 1), BRM_WAVE);
    const bool cbad = (__ballot(act && (col < 0 || col >= a.N || (gl > p_off && prev >= col))) & gmask) != 0ull;
    const bool run = ok && !cbad;
    const ShortMerged m = short_merge<true, G>(col, p, val, act && run, run ? ipn : 0, gl, gbase, gmask);
    const int nout = run ? m.nout : 0;
    if (gl == 0) {
        rcnt[slot] = nout;
        rip[slot] = run ? ipn : 0;
        if (bad || cbad) atomicOr(&st->err, (int)DERR_INVALID);
        if (miss && has) atomicOr(&st->miss, 1);
    }
    __syncthreads();
    // ---- the workgroup's rows in C: a scan over its rows, look-back over
    // the workgroups before it
    if (w == 0) {
        const int64_t c = lane < RPB ? rcnt[lane] : 0;
        const int64_t inc = wave_incl_scan(c);
        const int64_t agg = wave_last(inc);
        const int64_t ips = wave_sum(lane < RPB ? (int64_t)rip[lane] : (int64_t)0);
        const int64_t base = look_back(lb, b, agg);
        if (lane < RPB) rbase[lane] = base + inc - c;
        // nnz(C) and the products: one atomic each per workgroup, spread over
        // SMALL_SLOTS counters so that no address is hot
        if (lane == 0 && agg) atomicAdd((unsigned long long*)&st->nnz[b % SMALL_SLOTS], (unsigned long long)agg);
        if (lane == 1 && ips) atomicAdd((unsigned long long*)&st->ip[b % SMALL_SLOTS], (unsigned long long)ips);
    }
    __syncthreads();
    if (has) {
        const int64_t cb = rbase[slot];
        if (gl == 0) {
            crpt[row] = cb;
            if (row == a.M - 1) crpt[a.M] = cb + nout;
        }
        if (run && m.alive) {
            ccol[cb + m.idx] = m.col;
            cval[cb + m.idx] = m.val;
        }
    }
}

template <int G>
int64_t row_blocks_of(int M) {
    constexpr int RPB = SM_WPB * (BRM_WAVE / G);
    return ((int64_t)M + RPB - 1) / RPB;
}

}  // namespace

int64_t small_state_words(int M, int group) {
    const int64_t nb = group == 32 ? row_blocks_of<32>(M) : row_blocks_of<64>(M);
    return (int64_t)(sizeof(SmallState) / 8) + nb;
}

hipError_t launch_small(const RowArgs& a, int64_t nnzA, int64_t nnzB, int64_t* crpt, int* ccol, double* cval,
                        SmallState* st, int group, hipStream_t s) {
    hipError_t e = launch_fill_i64((int64_t*)st, small_state_words(a.M, group), 0, s);
    if (e != hipSuccess) return e;
    uint64_t* lb = reinterpret_cast<uint64_t*>(st) + sizeof(SmallState) / 8;
    const int64_t rb = group == 32 ? row_blocks_of<32>(a.M) : row_blocks_of<64>(a.M);
    const int64_t bb = ((int64_t)a.K + 1 + SM_WPB - 1) / SM_WPB;
    if (rb + bb > INT32_MAX) return hipErrorInvalidValue;
    {
        ProfScope ps_("k_small", s);
        if (group == 32)
            k_small<32><<<(unsigned)(rb + bb), SM_NT, 0, s>>>(a, nnzA, nnzB, (int)rb, crpt, ccol, cval, st, lb);
        else
            k_small<64><<<(unsigned)(rb + bb), SM_NT, 0, s>>>(a, nnzA, nnzB, (int)rb, crpt, ccol, cval, st, lb);
    }
    return hipGetLastError();
}

}  // namespace brm
