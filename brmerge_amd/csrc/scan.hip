// Device-wide exclusive prefix sums (int64 output), reduce-then-scan.
//
// Replaces opt_prefix_sum / para_prefix_sum (BRMerge-SpGEMM/inc/spgemm.h:147-187)
// and fuses the per-A-entry product count of compute_flop (spgemm.h:39-41):
// scan_alen() turns each A entry j into alen[j] = |B row acol[j]| on the fly,
// so out[j] is the global index of entry j's first intermediate product and
// out[arpt[i+1]] - out[arpt[i]] is row i's product count.
//
// Each block owns SCAN_TILE consecutive elements; within the block a
// wavefront scan (shuffles) plus an LDS pass over the 4 wave totals scans 256
// elements per step.  Block totals are scanned by a single block.
#include "brm_device.h"
#include "brm_internal.h"

namespace brm {
namespace {

constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int64_t SCAN_TILE = SCAN_NT * SCAN_ITEMS;

// A column outside [0, K) counts 0 products (validation reports it), so no
// B row pointer outside the array is read.
struct AlenF {
    const int* acol;
    const int* brpt;
    int K;
    __device__ int64_t operator()(int64_t j) const {
        const int c = acol[j];
        if ((unsigned)c >= (unsigned)K) return 0;
        return (int64_t)(brpt[c + 1] - brpt[c]);
    }
};
struct I64F {
    const int64_t* p;
    __device__ int64_t operator()(int64_t j) const { return p[j]; }
};
struct I32F {
    const int* p;
    __device__ int64_t operator()(int64_t j) const { return (int64_t)p[j]; }
};

template <typename F>
__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(F f, int64_t n, int64_t* part) {
    __shared__ int64_t sm[SCAN_NT / BRM_WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t acc = 0;
#pragma unroll 4
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + it * SCAN_NT + threadIdx.x;
        if (idx < n) acc += f(idx);
    }
    int64_t tot = block_sum<SCAN_NT>(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Exclusive scan of part[0..nb) in place; part[nb] = total.  One block.
__global__ __launch_bounds__(SCAN_NT) void k_scan_partials(int64_t* part, int64_t nb) {
    __shared__ int64_t sm[SCAN_NT / BRM_WAVE + 1];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN_NT) {
        int64_t idx = base + threadIdx.x;
        int64_t x = idx < nb ? part[idx] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan<SCAN_NT>(x, sm, &tot);
        if (idx < nb) part[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

template <typename F>
__global__ __launch_bounds__(SCAN_NT) void k_scan_down(F f, int64_t n, const int64_t* part, int64_t nb,
                                                       int64_t* out) {
    __shared__ int64_t sm[SCAN_NT / BRM_WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t carry = part[blockIdx.x];
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + it * SCAN_NT + threadIdx.x;
        if (base + it * SCAN_NT >= n) break;  // uniform across the block
        int64_t x = idx < n ? f(idx) : 0;
        int64_t tot;
        int64_t ex = block_excl_scan<SCAN_NT>(x, sm, &tot);
        if (idx < n) out[idx] = carry + ex;
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = part[nb];
}

// Down pass that also forms its own block's carry: the sum of the tile
// totals part[0 .. blockIdx.x), reduced by the block (for up to
// SCAN_FUSE_TILES tiles, so that this costs a few loads per thread).  Saves
// the one-block scan of the tile totals and its launch.
constexpr int64_t SCAN_FUSE_TILES = 8192;
template <typename F>
__global__ __launch_bounds__(SCAN_NT) void k_scan_down_fused(F f, int64_t n, const int64_t* part, int64_t* out) {
    __shared__ int64_t sm[SCAN_NT / BRM_WAVE + 1];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t pre = 0;
    for (int64_t b = threadIdx.x; b < blockIdx.x; b += SCAN_NT) pre += part[b];
    int64_t carry = block_sum<SCAN_NT>(pre, sm);
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + it * SCAN_NT + threadIdx.x;
        if (base + it * SCAN_NT >= n) break;  // uniform across the block
        int64_t x = idx < n ? f(idx) : 0;
        int64_t tot;
        int64_t ex = block_excl_scan<SCAN_NT>(x, sm, &tot);
        if (idx < n) out[idx] = carry + ex;
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = carry;
}

__global__ void k_zero1(int64_t* out) { out[0] = 0; }

template <typename F>
hipError_t scan_impl(F f, int64_t n, int64_t* out, hipStream_t s) {
    if (n <= 0) {
        {
            ProfScope ps_("k_zero1", s);
            k_zero1<<<1, 1, 0, s>>>(out);
        }
        return hipGetLastError();
    }
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t* part = (int64_t*)dev_alloc(sizeof(int64_t) * (size_t)(nb + 1));
    if (!part) return hipErrorOutOfMemory;
    {
        ProfScope ps_("k_scan_reduce", s);
        k_scan_reduce<F><<<(unsigned)nb, SCAN_NT, 0, s>>>(f, n, part);
    }
    if (nb <= SCAN_FUSE_TILES) {
        ProfScope ps_("k_scan_down", s);
        k_scan_down_fused<F><<<(unsigned)nb, SCAN_NT, 0, s>>>(f, n, part, out);
    } else {
        {
            ProfScope ps_("k_scan_partials", s);
            k_scan_partials<<<1, SCAN_NT, 0, s>>>(part, nb);
        }
        ProfScope ps_("k_scan_down", s);
        k_scan_down<F><<<(unsigned)nb, SCAN_NT, 0, s>>>(f, n, part, nb, out);
    }
    hipError_t e = hipGetLastError();
    dev_defer_free(part);  // still referenced by the queued kernels
    return e;
}

}  // namespace

hipError_t scan_alen(const int* acol, const int* brpt, int K, int64_t n, int64_t* out, hipStream_t s) {
    return scan_impl(AlenF{acol, brpt, K}, n, out, s);
}
hipError_t scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s) {
    return scan_impl(I64F{in}, n, out, s);
}
hipError_t scan_i32(const int* in, int64_t n, int64_t* out, hipStream_t s) {
    return scan_impl(I32F{in}, n, out, s);
}

}  // namespace brm
