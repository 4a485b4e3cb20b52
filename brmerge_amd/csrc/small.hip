// Short-row products in two launches and one host round trip.
//
// A product whose rows are all short (at most G = 64, or 32, lists and
// products) spends more on the general pipeline's dozen launches and two
// host round trips than on its kernels.  Such a product runs as:
//   1. one fill clearing the state: flags, counters, workgroup tickets and
//      look-back states;
//   2. k_small<G>: each workgroup takes a ticket (tickets are handed out in
//      the order workgroups start) and with it a run of consecutive rows,
//      in P passes of one row per group of G lanes.  The passes are
//      batched stage by stage -- all row pointers, then all list
//      descriptors, then all B entries -- so that each dependent load is
//      issued for every pass before the first one is waited on.  Each group
//      checks its rows' own indices, finds the list offsets by a group scan
//      (no global product scan) and merges the row in registers with the
//      short tier's code (short_merge.h).  The workgroup's nnz is scanned
//      across workgroups by decoupled look-back (a workgroup waits only on
//      workgroups that started before it), and every row is written straight
//      from registers into C at its place, with C's row pointer.
//      Workgroups past the row workgroups check all of B, one B row per
//      thread (row pointers, column range, sorted rows), as the general
//      validation does.
// C is allocated beforehand for G entries per row, a bound of nnz(C) when
// every row is short.  The host reads the state once, at the end: a row with
// more than G lists or products, or a B row with more than 64 entries, sets a
// flag, and the product then goes through the general pipeline instead.
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"
#include "short_merge.h"

namespace brm {
namespace {

// Look-back states: bits 62-63 the state (1: aggregate, 2: inclusive
// prefix), bits 0-61 the value.  Each state word carries its own value, so
// relaxed atomics (no cache invalidation or write-back) are enough.
constexpr uint64_t LB_AGG = 1ull << 62, LB_INC = 2ull << 62, LB_VAL = (1ull << 62) - 1ull;

// Exclusive prefix of this workgroup's `agg` over the workgroups before
// ticket b; called by all of wavefront 0.  Lane l reads the state l + 1
// below the window's top, 64 states per step.
__device__ int64_t look_back(uint64_t* lb, int b, int64_t agg) {
    const int lane = lane_id();
    int64_t prefix = 0;
    if (b > 0) {
        if (lane == 0) __hip_atomic_store(&lb[b], LB_AGG | (uint64_t)agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int hi = b - 1;;) {
            const int q = hi - lane;
            const uint64_t v = q >= 0 ? __hip_atomic_load(&lb[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : LB_INC;  // before workgroup 0: an empty prefix
            const uint64_t incm = __ballot((v & ~LB_VAL) == LB_INC);
            const int first = incm ? __ffsll((long long)incm) - 1 : BRM_WAVE;  // nearest prefix
            const uint64_t upto = first >= BRM_WAVE - 1 ? ~0ull : ((2ull << first) - 1ull);
            if ((__ballot(v == 0) & upto) != 0ull) {  // a state not yet published
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += wave_sum((lane <= first) ? (int64_t)(v & LB_VAL) : (int64_t)0);
            if (first < BRM_WAVE) break;
            hi -= BRM_WAVE;
        }
    }
    if (lane == 0)
        __hip_atomic_store(&lb[b], LB_INC | (uint64_t)(prefix + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

// Workgroup shape per group size: threads and passes.  A workgroup takes
// RPP * P consecutive rows (RPP rows per pass); wavefront 0 scans them, R per
// lane.  More rows per workgroup means fewer look-back states to chain.
constexpr int SM_NT32 = 512, SM_P32 = 4;   // G = 32
constexpr int SM_NT64 = 1024, SM_P64 = 4;  // G = 64
// B rows longer than this are left to the general pipeline's validation.
constexpr int SM_BROW_MAX = 64;

template <int G>
struct SmallShape {
    static constexpr int NT = G == 32 ? SM_NT32 : SM_NT64;  // threads
    static constexpr int P = G == 32 ? SM_P32 : SM_P64;     // passes
    static constexpr int WPB = NT / BRM_WAVE;                // wavefronts
    static constexpr int RPW = BRM_WAVE / G;                 // rows per wavefront and pass
    static constexpr int RPP = WPB * RPW;                    // rows per pass
    static constexpr int ROWS = RPP * P;                     // rows per workgroup
    static constexpr int R = ROWS / BRM_WAVE;                // rows per lane of the scan
    static_assert(R * BRM_WAVE == ROWS && R >= 1, "row layout");
};

template <int G>
__global__ __launch_bounds__(SmallShape<G>::NT) void k_small(RowArgs a, int64_t nnzA, int64_t nnzB, int row_blocks,
                                                              int64_t* __restrict__ crpt, int* __restrict__ ccol,
                                                              double* __restrict__ cval, SmallState* __restrict__ st,
                                                              uint64_t* __restrict__ lb) {
    using S = SmallShape<G>;
    constexpr int P = S::P;
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / BRM_WAVE);
    if ((int)blockIdx.x >= row_blocks) {
        // ---- B: one row per thread; entries 8 at a time (indices clamped
        // into the row, so that every load is in bounds)
        const int r = ((int)blockIdx.x - row_blocks) * S::NT + (int)threadIdx.x;
        bool bad = false, lng = false;
        if (r < a.K) {
            const int b0 = a.brpt[r], b1 = a.brpt[r + 1];
            if (b0 < 0 || b0 > b1 || (int64_t)b1 > nnzB) {
                bad = true;
            } else if (b1 - b0 > SM_BROW_MAX) {
                lng = true;
            } else {
                int prev = -1;
                for (int k0 = b0; k0 < b1; k0 += 8) {
                    int c[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) c[j] = a.bcol[min(k0 + j, b1 - 1)];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (k0 + j < b1) {
                            bad |= c[j] <= prev || c[j] >= a.N;
                            prev = c[j];
                        }
                    }
                }
            }
        } else if (r == a.K) {
            bad = a.brpt[0] != 0 || (int64_t)a.brpt[a.K] != nnzB;
        }
        if (__ballot(bad) != 0ull && lane == 0) atomicOr(&st->err, (int)DERR_INVALID);
        if (__ballot(lng) != 0ull && lane == 0) atomicOr(&st->miss, 1);
        return;
    }
    __shared__ int ticket;
    __shared__ int rcnt[S::ROWS];      // nnz of the workgroup's rows
    __shared__ int rip[S::ROWS];       // their products
    __shared__ int64_t rbase[S::ROWS];  // their C offsets
    if (threadIdx.x == 0) ticket = atomicAdd(&st->ticket, 1);
    __syncthreads();
    const int b = ticket;
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint64_t gmask = G == BRM_WAVE ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const int slot = w * S::RPW + lane / G;  // the group's row within a pass
    bool bad = false, miss = false;          // group-uniform
    // ---- stage 1: row pointers (all passes' loads first), checked before
    // they are followed; the checks use | and & so that no branch separates
    // one pass's loads from the next
    int a0[P], a1[P], L[P];
    bool ok[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int row = min(b * S::ROWS + p * S::RPP + slot, a.M - 1);
        a0[p] = a.arpt[row];
        a1[p] = a.arpt[row + 1];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int row = b * S::ROWS + p * S::RPP + slot;
        const bool has = row < a.M;
        const bool rb = has & ((a0[p] < 0) | (a0[p] > a1[p]) | ((int64_t)a1[p] > nnzA) | ((row == 0) & (a0[p] != 0)) |
                               ((row == a.M - 1) & ((int64_t)a1[p] != nnzA)));
        const bool ms = has & !rb & (a1[p] - a0[p] > G);
        bad |= rb;
        miss |= ms;
        ok[p] = has & !rb & !ms;
        L[p] = ok[p] ? a1[p] - a0[p] : 0;
    }
    // ---- stage 2: list descriptors (A column and value, B row bounds)
    int acl[P];
    double lav[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int idx = gl < L[p] ? a0[p] + gl : 0;  // nnz(A) >= 1
        acl[p] = a.acol[idx];
        lav[p] = a.aval[idx];
    }
    int lbst[P], lend[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const bool cok = (gl < L[p]) & (acl[p] >= 0) & (acl[p] < a.K);
        const int cc = cok ? acl[p] : 0;  // K >= 1
        lbst[p] = a.brpt[cc];
        lend[p] = a.brpt[cc + 1];
    }
    // ---- stage 3: list offsets, each lane's product, its B entry
    int pl[P], bcl[P], ipn[P];
    double pav[P], bvl[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const bool v = gl < L[p];
        const bool lbad = v & ((acl[p] < 0) | (acl[p] >= a.K) | (lbst[p] < 0) | (lbst[p] > lend[p]) |
                               ((int64_t)lend[p] > nnzB));
        const bool gbad = (__ballot(lbad) & gmask) != 0ull;
        bad |= gbad;
        // lengths capped at G + 1: a longer row is a miss either way, and the
        // group's sum stays within int
        const int bl = (v & !gbad) ? min(lend[p] - lbst[p], G + 1) : 0;
        int incl = wave_incl_scan(bl);
        if (G < BRM_WAVE) incl -= __shfl(incl, (gbase + BRM_WAVE - 1) & (BRM_WAVE - 1), BRM_WAVE) * (gbase > 0);
        const int ip = __shfl(incl, gbase + G - 1, BRM_WAVE);
        const bool ms = ok[p] & !gbad & (ip > G);
        miss |= ms;
        const bool run = ok[p] & !gbad & !ms;
        ipn[p] = run ? ip : 0;
        const int off = v ? incl - bl : INT_MAX;
        // product lane -> its list (largest q with off[q] <= gl), as the short tier
        int q = 0;
#pragma unroll
        for (int step = G / 2; step >= 1; step >>= 1) {
            const int cand = q + step;
            const int oc = __shfl(off, gbase + (cand & (G - 1)), BRM_WAVE);
            if (cand < L[p] && oc <= gl) q = cand;
        }
        const int src = gbase + (q & (G - 1));
        const int q_off = __shfl(off, src, BRM_WAVE);
        const int q_bst = __shfl(lbst[p], src, BRM_WAVE);
        pav[p] = __shfl(lav[p], src, BRM_WAVE);
        pl[p] = q;
        const int bi = gl < ipn[p] ? q_bst + (gl - q_off) : 0;  // nnz(B) >= 1
        bcl[p] = a.bcol[bi];
        bvl[p] = a.bval[bi];
    }
    // ---- stage 4: the merges
    const bool narrow = short_keys_fit<G>(a.N);
    ShortMerged m[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const bool act = gl < ipn[p];
        m[p] = short_merge<true, G>(act ? bcl[p] : INT_MAX, pl[p], act ? fmul(pav[p], bvl[p]) : 0.0, act, ipn[p], gl,
                                    gbase, gmask, narrow);
        if (gl == 0) {
            rcnt[p * S::RPP + slot] = ipn[p] > 0 ? m[p].nout : 0;
            rip[p * S::RPP + slot] = ipn[p];
        }
    }
    if (__ballot(bad) != 0ull && lane == 0) atomicOr(&st->err, (int)DERR_INVALID);
    if (__ballot(miss) != 0ull && lane == 0) atomicOr(&st->miss, 1);
    __syncthreads();
    // ---- the workgroup's rows in C: a scan over its rows, look-back over
    // the workgroups before it; C's row pointer from wavefront 0
    if (w == 0) {
        int64_t local = 0, ips = 0;
#pragma unroll
        for (int j = 0; j < S::R; ++j) {
            local += rcnt[lane * S::R + j];
            ips += rip[lane * S::R + j];
        }
        const int64_t inc = wave_incl_scan(local);
        const int64_t agg = wave_last(inc);
        ips = wave_sum(ips);
        const int64_t base = look_back(lb, b, agg);
        int64_t off = base + inc - local;
#pragma unroll
        for (int j = 0; j < S::R; ++j) {
            const int k = lane * S::R + j;
            const int row = b * S::ROWS + k;
            rbase[k] = off;
            if (row < a.M) crpt[row] = off;
            off += rcnt[k];
            if (row == a.M - 1) crpt[a.M] = off;
        }
        // nnz(C) and the products: one atomic each per workgroup, spread over
        // SMALL_SLOTS counters so that no address is hot
        if (lane == 0 && agg) atomicAdd((unsigned long long*)&st->nnz[b % SMALL_SLOTS], (unsigned long long)agg);
        if (lane == 1 && ips) atomicAdd((unsigned long long*)&st->ip[b % SMALL_SLOTS], (unsigned long long)ips);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (ipn[p] > 0 && m[p].alive) {
            const int64_t cb = rbase[p * S::RPP + slot];
            ccol[cb + m[p].idx] = m[p].col;
            cval[cb + m[p].idx] = m[p].val;
        }
    }
}

template <int G>
int64_t row_blocks_of(int M) {
    return ((int64_t)M + SmallShape<G>::ROWS - 1) / SmallShape<G>::ROWS;
}

template <int G>
hipError_t launch_small_g(const RowArgs& a, int64_t nnzA, int64_t nnzB, int64_t* crpt, int* ccol, double* cval,
                          SmallState* st, uint64_t* lb, hipStream_t s) {
    using S = SmallShape<G>;
    const int64_t rb = row_blocks_of<G>(a.M);
    const int64_t bb = ((int64_t)a.K + 1 + S::NT - 1) / S::NT;
    if (rb + bb > INT32_MAX) return hipErrorInvalidValue;
    ProfScope ps_("k_small", s);
    k_small<G><<<(unsigned)(rb + bb), S::NT, 0, s>>>(a, nnzA, nnzB, (int)rb, crpt, ccol, cval, st, lb);
    return hipGetLastError();
}

}  // namespace

int64_t small_state_words(int M, int group) {
    const int64_t nb = group == 32 ? row_blocks_of<32>(M) : row_blocks_of<64>(M);
    return (int64_t)(sizeof(SmallState) / 8) + nb;
}

hipError_t launch_small(const RowArgs& a, int64_t nnzA, int64_t nnzB, int64_t* crpt, int* ccol, double* cval,
                        SmallState* st, int group, hipStream_t s) {
    if (a.M < 1 || a.K < 1 || nnzA < 1 || nnzB < 1) return hipErrorInvalidValue;
    hipError_t e = launch_fill_i64((int64_t*)st, small_state_words(a.M, group), 0, s);
    if (e != hipSuccess) return e;
    uint64_t* lb = reinterpret_cast<uint64_t*>(st) + sizeof(SmallState) / 8;
    return group == 32 ? launch_small_g<32>(a, nnzA, nnzB, crpt, ccol, cval, st, lb, s)
                       : launch_small_g<64>(a, nnzA, nnzB, crpt, ccol, cval, st, lb, s);
}

}  // namespace brm
