// CSR assembly helpers.
//
// Upper compaction: rows merged into C-bar at their product offsets
// (ipA[arpt[row]], the reference's scan_floprC, spgemm.h:399-400) are packed
// into C at crpt[row] (spgemm.h:427-445).  Each workgroup owns a contiguous
// range of C, locates the row of its first element by binary search on crpt
// and walks rows forward, so reads and writes are both coalesced.
#include <algorithm>

#include "brm_device.h"
#include "kernels.h"

namespace brm {
namespace {

constexpr int CP_NT = 256;
constexpr int CP_ITEMS = 16;
constexpr int64_t CP_TILE = CP_NT * CP_ITEMS;

__global__ __launch_bounds__(CP_NT) void k_compact(const int64_t* __restrict__ crpt, int M, RowArgs a,
                                                   const int* __restrict__ cbar_col,
                                                   const double* __restrict__ cbar_val, int* __restrict__ ccol,
                                                   double* __restrict__ cval, int64_t nnzC) {
    __shared__ int r0s;
    const int64_t q0 = (int64_t)blockIdx.x * CP_TILE;
    const int64_t q1 = min(nnzC, q0 + CP_TILE);
    if (threadIdx.x == 0) r0s = upper_idx(crpt, 0, M, q0);
    __syncthreads();
    // a thread's outputs ascend: its row's C-bar base and end stay in
    // registers until an output crosses into a later row
    int r = r0s;
    int64_t sbase = 0, rend = -1;
    for (int64_t q = q0 + threadIdx.x; q < q1; q += CP_NT) {
        if (q >= rend) {
            r = gallop_idx(crpt, r, M, q);
            sbase = a.ipA[a.arpt[r]] - crpt[r];
            rend = crpt[r + 1];
        }
        const int64_t src = sbase + q;
        ccol[q] = cbar_col[src];
        cval[q] = cbar_val[src];
    }
}

// Plain streaming copy (16-byte vectors), used to calibrate bandwidth counters.
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// Order-independent 64-bit checksum of 4-byte words: the sum (mod 2^64) of
// a SplitMix64 finalisation of (word index, word).  Two arrays that differ
// in any word differ in the sum except with probability about 2^-64.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_checksum(const uint32_t* __restrict__ w, int64_t n, uint64_t* out) {
    uint64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        acc += mix64((uint64_t)i * 0x9E3779B97F4A7C15ull + w[i]);
    acc = wave_sum(acc);
    if (lane_id() == 0) atomicAdd((unsigned long long*)out, (unsigned long long)acc);
}

__global__ void k_rpt_to_i32(const int64_t* __restrict__ in, int64_t n, int* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (int)in[i];
}

__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int* p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_compact(const int64_t* crpt, int M, const RowArgs& a, const int* cbar_col, const double* cbar_val,
                          int* ccol, double* cval, int64_t nnzC, hipStream_t s) {
    if (nnzC <= 0) return hipSuccess;
    {
        ProfScope ps_("k_compact", s);
        k_compact<<<nblk(nnzC, (int)CP_TILE), CP_NT, 0, s>>>(crpt, M, a, cbar_col, cbar_val, ccol, cval, nnzC);
    }
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 <= 0) return hipSuccess;
    {
        ProfScope ps_("k_copy", s);
        k_copy<<<2048, 256, 0, s>>>((const uint4*)src, (uint4*)dst, n16);
    }
    return hipGetLastError();
}

hipError_t launch_checksum(const void* p, int64_t nwords, uint64_t* out, hipStream_t s) {
    hipError_t e = launch_fill_i64((int64_t*)out, 1, 0, s);
    if (e != hipSuccess || nwords <= 0) return e;
    {
        ProfScope ps_("k_checksum", s);
        k_checksum<<<(unsigned)std::min<int64_t>(nblk(nwords, 256), 8192), 256, 0, s>>>((const uint32_t*)p, nwords,
                                                                                      out);
    }
    return hipGetLastError();
}

hipError_t launch_rpt_to_i32(const int64_t* in, int64_t n, int* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_rpt_to_i32", s);
        k_rpt_to_i32<<<nblk(n, 256), 256, 0, s>>>(in, n, out);
    }
    return hipGetLastError();
}

// Fills use kernels rather than hipMemsetAsync so that the sequence stays
// valid inside captured HIP graphs.
hipError_t launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_fill_i64", s);
        k_fill_i64<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
    }
    return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int64_t n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    {
        ProfScope ps_("k_fill_i32", s);
        k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
    }
    return hipGetLastError();
}

}  // namespace brm
