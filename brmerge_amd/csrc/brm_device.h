// Device-side building blocks shared by the BRMerge kernels (gfx950, wave64).
//
// Everything here is plain HIP C++ for CDNA4: 64-lane wavefronts, DPP lane
// moves for wave scans, LDS for block scans.  Value arithmetic uses explicit
// round-to-nearest multiply and add (no contraction into FMA), because the
// reference stores each product aval*bval before any add
// (BRMerge-SpGEMM/inc/spgemm.h:621,653).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define BRM_WAVE 64

namespace brm {

__device__ __forceinline__ double fmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double fadd(double a, double b) { return __dadd_rn(a, b); }

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// ---------------------------------------------------------------- wave scans
// DPP lane moves (no LDS round trips).  Lane i takes the value of lane i - s
// of its row of 16 lanes (row_shr:s), or lane 15 / lane 31 (row_bcast15 /
// row_bcast31, written only into the rows of ROWS); lanes without a source
// take `id`.  Callers run with the whole wavefront active.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_mov(int x, int id) {
    return __builtin_amdgcn_update_dpp(id, x, CTRL, ROWS, 0xF, false);
}

template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T dpp_take(T x) {
    if constexpr (sizeof(T) == 4) {
        return (T)dpp_mov<CTRL, ROWS>((int)x, 0);
    } else {
        const uint64_t u = (uint64_t)x;
        const uint32_t lo = (uint32_t)dpp_mov<CTRL, ROWS>((int)(uint32_t)u, 0);
        const uint32_t hi = (uint32_t)dpp_mov<CTRL, ROWS>((int)(uint32_t)(u >> 32), 0);
        return (T)(((uint64_t)hi << 32) | lo);
    }
}

// Inclusive sum over the lanes: Hillis-Steele within rows of 16, then the
// row totals are carried by the two row broadcasts.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
    static_assert(std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8), "integer scans");
    x += dpp_take<0x111, 0xF>(x);  // row_shr:1
    x += dpp_take<0x112, 0xF>(x);  // row_shr:2
    x += dpp_take<0x114, 0xF>(x);  // row_shr:4
    x += dpp_take<0x118, 0xF>(x);  // row_shr:8
    x += dpp_take<0x142, 0xA>(x);  // row_bcast15 into rows 1 and 3
    x += dpp_take<0x143, 0xC>(x);  // row_bcast31 into rows 2 and 3
    return x;
}

// Value of the last lane, in every lane.
template <typename T>
__device__ __forceinline__ T wave_last(T x) {
    if constexpr (sizeof(T) == 4) {
        return (T)__builtin_amdgcn_readlane((int)x, BRM_WAVE - 1);
    } else {
        const uint64_t u = (uint64_t)x;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, BRM_WAVE - 1);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), BRM_WAVE - 1);
        return (T)(((uint64_t)hi << 32) | lo);
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
    return wave_last(wave_incl_scan(x));
}

// ---------------------------------------------------------------- block scans
// Exclusive scan over the block; returns this thread's exclusive prefix and
// the block total in *total.  `sm` must hold NT/64 + 1 elements of T.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* sm, T* total) {
    constexpr int NW = NT / BRM_WAVE;
    const int lane = lane_id();
    const int w = threadIdx.x / BRM_WAVE;
    T inc = wave_incl_scan(x);
    if (lane == BRM_WAVE - 1) sm[w] = inc;
    __syncthreads();
    if (threadIdx.x < BRM_WAVE) {
        T v = (lane < NW) ? sm[lane] : T(0);
        T vi = wave_incl_scan(v);
        if (lane < NW) sm[lane] = vi - v;
        if (lane == NW - 1) sm[NW] = vi;
    }
    __syncthreads();
    T res = sm[w] + inc - x;
    *total = sm[NW];
    __syncthreads();
    return res;
}

// Exclusive scan with a single barrier.  `sm` (NT/64 elements) must not be
// written again before every thread of the block has passed another barrier.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan_1b(T x, T* sm, T* total) {
    constexpr int NW = NT / BRM_WAVE;
    const int lane = lane_id();
    const int w = threadIdx.x / BRM_WAVE;
    const T inc = wave_incl_scan(x);
    if (lane == BRM_WAVE - 1) sm[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const T v = sm[i];
        pre += (i < w) ? v : T(0);
        tot += v;
    }
    *total = tot;
    return pre + inc - x;
}

template <int NT, typename T>
__device__ __forceinline__ T block_sum(T x, T* sm) {
    T tot;
    (void)block_excl_scan<NT, T>(x, sm, &tot);
    return tot;
}

// ---------------------------------------------------------------- searches
// Largest k in [lo, hi) with a[k] <= key (a nondecreasing, a[lo] <= key).
template <typename T, typename K>
__device__ __forceinline__ int upper_idx(const T* a, int lo, int hi, K key) {
    // invariant: a[lo] <= key, answer in [lo, hi)
    int n = hi - lo;
    while (n > 1) {
        int half = n >> 1;
        int mid = lo + half;
        if ((K)a[mid] <= key) { lo = mid; n -= half; } else { n = half; }
    }
    return lo;
}

// Galloping variant: start from hint (a[hint] <= key), search forward.
template <typename T, typename K>
__device__ __forceinline__ int gallop_idx(const T* a, int hint, int hi, K key) {
    int lo = hint;
    int step = 1;
    int probe = lo + step;
    while (probe < hi && (K)a[probe] <= key) {
        lo = probe;
        step <<= 1;
        probe = lo + step;
    }
    int end = probe < hi ? probe : hi;
    return upper_idx(a, lo, end, key);
}

// Merge-path split of the stable (left-first on ties) merge of two sorted,
// duplicate-free column lists L[0..nL) and R[0..nR) at diagonal d: returns
// i (taken from L); j = d - i.  If L[i-1] == R[j] the pair would straddle the
// cut, so j is advanced by one and both copies fall in the earlier part.
template <typename CL, typename CR>
__device__ __forceinline__ void merge_split(CL colL, int nL, CR colR, int nR, int d, int* oi, int* oj) {
    int lo = d - nR; if (lo < 0) lo = 0;
    int hi = d < nL ? d : nL;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (colL(mid) <= colR(d - 1 - mid)) lo = mid + 1; else hi = mid;
    }
    int i = lo, j = d - lo;
    if (i > 0 && j < nR && colL(i - 1) == colR(j)) j++;
    *oi = i;
    *oj = j;
}

template <typename CL, typename CR>
__device__ __forceinline__ void merge_split64(CL colL, int64_t nL, CR colR, int64_t nR, int64_t d,
                                              int64_t* oi, int64_t* oj) {
    int64_t lo = d - nR; if (lo < 0) lo = 0;
    int64_t hi = d < nL ? d : nL;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (colL(mid) <= colR(d - 1 - mid)) lo = mid + 1; else hi = mid;
    }
    int64_t i = lo, j = d - lo;
    if (i > 0 && j < nR && colL(i - 1) == colR(j)) j++;
    *oi = i;
    *oj = j;
}

}  // namespace brm
