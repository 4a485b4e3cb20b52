// Small products in three launches and one host round trip.
//
// A product of few rows (at most SMALL_MAX_ROWS) is launch-bound: the
// general pipeline's dozen launches and two host round trips cost more than
// its kernels.  When every row has at most 64 lists and 64 products (the
// short tier), the whole product runs as:
//   1. one fill clearing the flags, counters and look-back states;
//   2. k_small_rows: one wavefront per row validates the row's own indices
//      (and the extra waves validate all of B, as the general pipeline does),
//      computes its list offsets by a wave scan (no global product scan),
//      merges the row in registers exactly as the short tier does
//      (short_merge.h) into a 64-slot stash, and counts nnz(C) and the
//      products; the host reads the flags and nnz(C) once and allocates C;
//   3. k_small_emit: C's row pointer by a single-pass scan with decoupled
//      look-back (each workgroup takes a ticket, so a workgroup only waits on
//      workgroups that started before it), and the stash copied into C.
// A product with a longer row is left to the general pipeline.
#include <limits.h>

#include <algorithm>

#include "brm_device.h"
#include "kernels.h"
#include "short_merge.h"

namespace brm {
namespace {

constexpr int SM_NT = 256;
constexpr int SM_WPB = SM_NT / BRM_WAVE;
constexpr int SM_EMIT_ROWS = 64;  // rows per workgroup of k_small_emit

// k_small_rows: workgroups [0, row_blocks) take one row per wavefront;
// workgroups past them validate B (row pointers, columns in range, sorted
// rows), one B row per wavefront.
__global__ __launch_bounds__(SM_NT) void k_small_rows(RowArgs a, int64_t nnzA, int64_t nnzB, int row_blocks,
                                                      int* __restrict__ scol, double* __restrict__ sval,
                                                      int64_t* __restrict__ rowcnt, SmallState* __restrict__ st) {
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
    // ---- A: one row per wavefront, every check local to the row
    const int row = (int)blockIdx.x * SM_WPB + w;
    const bool has = row < a.M;
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
    if (L > BRM_WAVE) {
        miss = true;
        L = 0;
    }
    // list lanes: B row start and length (checked), A value
    int lbst = 0;
    int64_t blen = 0;
    double lav = 0.0;
    if (!bad && lane < L) {
        const int c = a.acol[a0 + lane];
        if (c < 0 || c >= a.K) {
            bad = true;
        } else {
            const int b0 = a.brpt[c], b1 = a.brpt[c + 1];
            if (b0 < 0 || b0 > b1 || (int64_t)b1 > nnzB) bad = true;
            else {
                lbst = b0;
                blen = b1 - b0;
            }
        }
        lav = a.aval[a0 + lane];
    }
    bad = __ballot(bad) != 0ull;
    if (bad) blen = 0;
    const int64_t incl = wave_incl_scan(blen);
    const int64_t ip = wave_last(incl);
    if (ip > BRM_WAVE) miss = true;
    const bool ok = has && !bad && !miss;
    const int off = (lane < L) ? (int)(incl - blen) : INT_MAX;
    const int ipn = ok ? (int)ip : 0;
    // product lane -> its list (largest p with off[p] <= lane), as the short tier
    int p = 0;
#pragma unroll
    for (int step = BRM_WAVE / 2; step >= 1; step >>= 1) {
        const int cand = p + step;
        const int oc = __shfl(off, cand & (BRM_WAVE - 1), BRM_WAVE);
        if (cand < L && oc <= lane) p = cand;
    }
    const int p_off = __shfl(off, p, BRM_WAVE);
    const int p_bst = __shfl(lbst, p, BRM_WAVE);
    const double p_av = __shfl(lav, p, BRM_WAVE);
    const bool act = lane < ipn;
    int col = INT_MAX;
    double val = 0.0;
    if (act) {
        const int bi = p_bst + (lane - p_off);
        col = a.bcol[bi];
        val = fmul(p_av, a.bval[bi]);
    }
    // columns in range and strictly increasing within each list
    const int prev = __shfl(col, (lane + BRM_WAVE - 1) & (BRM_WAVE - 1), BRM_WAVE);
    const bool cbad = __ballot(act && (col < 0 || col >= a.N || (lane > p_off && prev >= col))) != 0ull;
    const bool run = ok && !cbad;
    const ShortMerged m = short_merge<true, BRM_WAVE>(col, p, val, act && run, run ? ipn : 0, lane, 0, ~0ull);
    if (run && m.alive) {
        scol[(int64_t)row * BRM_WAVE + m.idx] = m.col;
        sval[(int64_t)row * BRM_WAVE + m.idx] = m.val;
    }
    // the workgroup's nnz(C) and products go to one of SMALL_SLOTS counters
    // (one atomic per workgroup, spread over slots: no single hot address)
    __shared__ int64_t part[2][SM_WPB];
    if (lane == 0) {
        if (has) rowcnt[row] = run ? m.nout : 0;
        if (bad || cbad) atomicOr(&st->err, (int)DERR_INVALID);
        if (miss && has) atomicOr(&st->miss, 1);
        part[0][w] = run ? m.nout : 0;
        part[1][w] = run ? ip : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t n = 0, q = 0;
#pragma unroll
        for (int i = 0; i < SM_WPB; ++i) {
            n += part[0][i];
            q += part[1][i];
        }
        const int sl = (int)blockIdx.x % SMALL_SLOTS;
        if (n) atomicAdd((unsigned long long*)&st->nnz[sl], (unsigned long long)n);
        if (q) atomicAdd((unsigned long long*)&st->ip[sl], (unsigned long long)q);
    }
}

// Look-back states: bits 62-63 the state (1: aggregate, 2: inclusive
// prefix), bits 0-61 the value.
constexpr uint64_t LB_AGG = 1ull << 62, LB_INC = 2ull << 62, LB_VAL = (1ull << 62) - 1ull;

__global__ __launch_bounds__(SM_NT) void k_small_emit(int M, const int64_t* __restrict__ rowcnt,
                                                      const int* __restrict__ scol, const double* __restrict__ sval,
                                                      int64_t* __restrict__ crpt, int* __restrict__ ccol,
                                                      double* __restrict__ cval, SmallState* __restrict__ st,
                                                      uint64_t* __restrict__ lb) {
    static_assert(SM_EMIT_ROWS == BRM_WAVE, "one row per lane of wavefront 0 in the scan");
    __shared__ int ticket;
    __shared__ int64_t prefix_s;
    __shared__ int rowoff[SM_EMIT_ROWS + 1];  // the workgroup's rows in its output range
    const int t = threadIdx.x;
    if (t == 0) ticket = atomicAdd(&st->ticket, 1);  // workgroups take rows in the order they start
    __syncthreads();
    const int b = ticket;
    const int r = b * SM_EMIT_ROWS + t;
    int64_t c = 0, ex = 0, agg = 0;
    if (t < BRM_WAVE) {
        c = r < M ? rowcnt[r] : 0;
        const int64_t inc = wave_incl_scan(c);
        ex = inc - c;
        agg = wave_last(inc);
    }
    if (t == 0) {
        int64_t prefix = 0;
        if (b == 0) {
            __hip_atomic_store(&lb[0], LB_INC | (uint64_t)agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&lb[b], LB_AGG | (uint64_t)agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            // every workgroup before b has started (tickets are taken in
            // order), so each state below is published eventually
            for (int q = b - 1; q >= 0;) {
                const uint64_t v = __hip_atomic_load(&lb[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (v == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += (int64_t)(v & LB_VAL);
                if ((v & ~LB_VAL) == LB_INC) break;
                --q;
            }
            __hip_atomic_store(&lb[b], LB_INC | (uint64_t)(prefix + agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        prefix_s = prefix;
    }
    __syncthreads();
    const int64_t base = prefix_s;
    if (t < BRM_WAVE) {
        rowoff[t] = (int)ex;
        if (t == BRM_WAVE - 1) rowoff[BRM_WAVE] = (int)agg;
        if (r < M) crpt[r] = base + ex;
        if (r == M - 1) crpt[M] = base + ex + c;
    }
    __syncthreads();
    // the workgroup's output range [base, base + agg) copied from the stash,
    // one entry per thread and step (coalesced stores, independent loads);
    // an entry's row is found by a search over the rows' offsets
    const int n = rowoff[SM_EMIT_ROWS];
    for (int q = t; q < n; q += SM_NT) {
        const int k = upper_idx(rowoff, 0, SM_EMIT_ROWS, q);
        const int64_t src = (int64_t)(b * SM_EMIT_ROWS + k) * BRM_WAVE + (q - rowoff[k]);
        ccol[base + q] = scol[src];
        cval[base + q] = sval[src];
    }
}

}  // namespace

// SmallState at st, then look-back states (one per k_small_emit workgroup),
// all cleared by one fill.
int64_t small_state_words(int M) { return (int64_t)(sizeof(SmallState) / 8) + (M + SM_EMIT_ROWS - 1) / SM_EMIT_ROWS; }

hipError_t launch_small(const RowArgs& a, int64_t nnzA, int64_t nnzB, int* scol, double* sval, int64_t* rowcnt,
                        SmallState* st, hipStream_t s) {
    hipError_t e = launch_fill_i64((int64_t*)st, small_state_words(a.M), 0, s);
    if (e != hipSuccess) return e;
    const int row_blocks = (a.M + SM_WPB - 1) / SM_WPB;
    const int b_blocks = (a.K + 1 + SM_WPB - 1) / SM_WPB;
    {
        ProfScope ps_("k_small_rows", s);
        k_small_rows<<<row_blocks + b_blocks, SM_NT, 0, s>>>(a, nnzA, nnzB, row_blocks, scol, sval, rowcnt, st);
    }
    return hipGetLastError();
}

hipError_t launch_small_emit(int M, const int64_t* rowcnt, const int* scol, const double* sval, int64_t* crpt,
                             int* ccol, double* cval, SmallState* st, hipStream_t s) {
    const int nb = (M + SM_EMIT_ROWS - 1) / SM_EMIT_ROWS;
    if (nb == 0) return hipSuccess;
    uint64_t* lb = reinterpret_cast<uint64_t*>(st) + sizeof(SmallState) / 8;
    {
        ProfScope ps_("k_small_emit", s);
        k_small_emit<<<nb, SM_NT, 0, s>>>(M, rowcnt, scol, sval, crpt, ccol, cval, st, lb);
    }
    return hipGetLastError();
}

}  // namespace brm
