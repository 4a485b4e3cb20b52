// CSR transpose on the device (the step before P^T * (A*P) in the AMG
// Galerkin product; the reference forms transposes on the host by sorting,
// Existing-SpGEMM/inc/convert.h:6-58).
//
// Entry e of row r becomes entry (col[e], r) of the transpose.  A stable
// radix sort of the entry indices by column (hipCUB) keeps the rows of every
// output row in increasing order, so the result is the canonical CSR of A^T;
// the output row pointer is the exclusive scan of the column counts.
#include <hipcub/hipcub.hpp>

#include "brm_device.h"
#include "pipeline.h"

namespace brm {
namespace {

constexpr int TP_NT = 256;

// row id of every entry; column counts
__global__ __launch_bounds__(TP_NT) void k_tp_rows(const int* __restrict__ rpt, int M, int* __restrict__ row_of,
                                                   const int* __restrict__ col, int64_t* __restrict__ cnt) {
    const int r = blockIdx.x * (TP_NT / BRM_WAVE) + threadIdx.x / BRM_WAVE;
    if (r >= M) return;
    for (int e = rpt[r] + lane_id(); e < rpt[r + 1]; e += BRM_WAVE) {
        row_of[e] = r;
        atomicAdd((unsigned long long*)&cnt[col[e]], 1ull);
    }
}

__global__ __launch_bounds__(TP_NT) void k_tp_gather(const int* __restrict__ perm, int64_t nnz,
                                                     const int* __restrict__ row_of, const double* __restrict__ val,
                                                     int* __restrict__ tcol, double* __restrict__ tval) {
    const int64_t k = (int64_t)blockIdx.x * TP_NT + threadIdx.x;
    if (k >= nnz) return;
    const int e = perm[k];
    tcol[k] = row_of[e];
    tval[k] = val[e];
}

__global__ __launch_bounds__(TP_NT) void k_tp_iota(int* __restrict__ p, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * TP_NT + threadIdx.x;
    if (k < n) p[k] = (int)k;
}

unsigned tp_blocks(int64_t n) { return (unsigned)((n + TP_NT - 1) / TP_NT); }

int transpose_dev(const int* rpt, const int* col, const double* val, int64_t nnz, int M, int N, int* trpt,
                  int* tcol, double* tval, hipStream_t s) {
    Scratch sc;
    int *row_of, *perm_in, *perm_out, *keys_out, *derr;
    int64_t *cnt, *tr64;
    BRMC(sc.alloc(&derr, 1, "error flag"));
    BRMC(sc.alloc(&row_of, nnz, "row ids"));
    BRMC(sc.alloc(&perm_in, nnz, "entry ids"));
    BRMC(sc.alloc(&perm_out, nnz, "sorted entry ids"));
    BRMC(sc.alloc(&keys_out, nnz, "sorted columns"));
    BRMC(sc.alloc(&cnt, N, "column counts"));
    BRMC(sc.alloc(&tr64, (int64_t)N + 1, "row pointer"));
    HIPC(launch_fill_i32(derr, 1, 0, s));
    // the input must be well formed: row pointers and column range
    RowArgs a{};
    a.arpt = rpt;
    a.acol = col;
    a.M = M;
    a.K = N;
    HIPC(launch_validate(a, nnz, -1, derr, s));
    BRMC(check_dev_err(derr, s, "transpose input"));
    HIPC(launch_fill_i64(cnt, N, 0, s));
    if (M > 0) {
        k_tp_rows<<<(unsigned)((M + TP_NT / BRM_WAVE - 1) / (TP_NT / BRM_WAVE)), TP_NT, 0, s>>>(rpt, M, row_of, col,
                                                                                             cnt);
        HIPC(hipGetLastError());
    }
    if (nnz > 0) {
        k_tp_iota<<<tp_blocks(nnz), TP_NT, 0, s>>>(perm_in, nnz);
        HIPC(hipGetLastError());
        int end_bit = 1;
        while (end_bit < 31 && ((int64_t)1 << end_bit) < N) ++end_bit;
        size_t tmp_bytes = 0;
        HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, col, keys_out, perm_in, perm_out, (int)nnz, 0,
                                                end_bit, s));
        unsigned char* tmp;
        BRMC(sc.alloc(&tmp, (int64_t)tmp_bytes, "sort scratch"));
        HIPC(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, col, keys_out, perm_in, perm_out, (int)nnz, 0,
                                                end_bit, s));
        k_tp_gather<<<tp_blocks(nnz), TP_NT, 0, s>>>(perm_out, nnz, row_of, val, tcol, tval);
        HIPC(hipGetLastError());
    }
    if (N > 0) {
        HIPC(scan_i64(cnt, N, tr64, s));
        HIPC(launch_rpt_to_i32(tr64, (int64_t)N + 1, trpt, s));
    } else {
        HIPC(launch_fill_i32(trpt, 1, 0, s));
    }
    HIPC(hipStreamSynchronize(s));
    return BRM_OK;
}

}  // namespace
}  // namespace brm

using namespace brm;

extern "C" int brm_csr_transpose_dev(const int* rpt, const int* col, const double* val, int64_t nnz, int M, int N,
                                     int* trpt, int* tcol, double* tval, void* stream) {
    if (M < 0 || N < 0 || nnz < 0 || nnz > INT32_MAX || !trpt || (M > 0 && !rpt) ||
        (nnz > 0 && (!col || !val || !tcol || !tval))) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    const int st = transpose_dev(rpt, col, val, nnz, M, N, trpt, tcol, tval, (hipStream_t)stream);
    dev_flush_deferred();
    return st;
}
