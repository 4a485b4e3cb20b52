// Vendor cross-check: the same product computed by rocSPARSE's generic
// SpGEMM (the analogue of the reference's MKL verification,
// BRMerge-SpGEMM/inc/mkl_spgemm.h:8-67 and src/brmerge_precise.cpp:86-93).
// Test and benchmark infrastructure only: libbrm_xcheck.so is a separate
// library and nothing in libbrmerge.so uses it.
//
// rocSPARSE sums in its own order, so its values agree with BRMerge's to
// rounding (the tests use the 1e-12 relative bound), and its structure must
// match exactly.  int32 indices: nnz(C) must fit int32.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>
#include <stdint.h>
#include <stdio.h>

#include <chrono>

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Guard {
    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr A = nullptr, B = nullptr, C = nullptr, D = nullptr;
    void* buf = nullptr;
    int* drpt = nullptr;
    ~Guard() {
        if (A) rocsparse_destroy_spmat_descr(A);
        if (B) rocsparse_destroy_spmat_descr(B);
        if (C) rocsparse_destroy_spmat_descr(C);
        if (D) rocsparse_destroy_spmat_descr(D);
        if (h) rocsparse_destroy_handle(h);
        if (buf) (void)hipFree(buf);
        if (drpt) (void)hipFree(drpt);
    }
};

}  // namespace

#define RS(x)                                                                              \
    do {                                                                                   \
        rocsparse_status s_ = (x);                                                         \
        if (s_ != rocsparse_status_success) {                                              \
            fprintf(stderr, "brm_xcheck: %s failed (%d)\n", #x, (int)s_);                  \
            return 4;                                                                      \
        }                                                                                  \
    } while (0)
#define HC(x)                                                                              \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "brm_xcheck: %s failed (%s)\n", #x, hipGetErrorString(e_));    \
            return 4;                                                                      \
        }                                                                                  \
    } while (0)

extern "C" {

// C = A * B by rocSPARSE.  All pointers are device pointers; *crpt, *ccol,
// *cval are allocated with hipMalloc (release with brm_xcheck_free).
// *total_s: the whole call (all stages and allocations); *compute_s: the
// compute stage alone.
int brm_xcheck_spgemm(const int* arpt, const int* acol, const double* aval, int64_t nnzA, const int* brpt,
                      const int* bcol, const double* bval, int64_t nnzB, int M, int K, int N, int** crpt, int** ccol,
                      double** cval, int64_t* cnnz, double* total_s, double* compute_s) {
    const double t0 = now_s();
    Guard g;
    RS(rocsparse_create_handle(&g.h));
    const rocsparse_indextype i32 = rocsparse_indextype_i32;
    const rocsparse_index_base z = rocsparse_index_base_zero;
    const rocsparse_datatype f64 = rocsparse_datatype_f64_r;
    RS(rocsparse_create_csr_descr(&g.A, M, K, nnzA, (void*)arpt, (void*)acol, (void*)aval, i32, i32, z, f64));
    RS(rocsparse_create_csr_descr(&g.B, K, N, nnzB, (void*)brpt, (void*)bcol, (void*)bval, i32, i32, z, f64));
    int* c_rpt = nullptr;
    HC(hipMalloc(&c_rpt, sizeof(int) * ((size_t)M + 1)));
    HC(hipMalloc(&g.drpt, sizeof(int) * ((size_t)M + 1)));
    HC(hipMemset(g.drpt, 0, sizeof(int) * ((size_t)M + 1)));
    RS(rocsparse_create_csr_descr(&g.C, M, N, 0, c_rpt, nullptr, nullptr, i32, i32, z, f64));
    RS(rocsparse_create_csr_descr(&g.D, M, N, 0, g.drpt, nullptr, nullptr, i32, i32, z, f64));
    const double alpha = 1.0, beta = 0.0;
    const rocsparse_operation none = rocsparse_operation_none;
    size_t bs = 0;
    RS(rocsparse_spgemm(g.h, none, none, &alpha, g.A, g.B, &beta, g.D, g.C, f64, rocsparse_spgemm_alg_default,
                        rocsparse_spgemm_stage_buffer_size, &bs, nullptr));
    HC(hipMalloc(&g.buf, bs > 0 ? bs : 16));
    RS(rocsparse_spgemm(g.h, none, none, &alpha, g.A, g.B, &beta, g.D, g.C, f64, rocsparse_spgemm_alg_default,
                        rocsparse_spgemm_stage_nnz, &bs, g.buf));
    int64_t rows = 0, cols = 0, nnz = 0;
    RS(rocsparse_spmat_get_size(g.C, &rows, &cols, &nnz));
    int* c_col = nullptr;
    double* c_val = nullptr;
    HC(hipMalloc(&c_col, sizeof(int) * (size_t)(nnz > 0 ? nnz : 1)));
    HC(hipMalloc(&c_val, sizeof(double) * (size_t)(nnz > 0 ? nnz : 1)));
    RS(rocsparse_csr_set_pointers(g.C, c_rpt, c_col, c_val));
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, nullptr));
    RS(rocsparse_spgemm(g.h, none, none, &alpha, g.A, g.B, &beta, g.D, g.C, f64, rocsparse_spgemm_alg_default,
                        rocsparse_spgemm_stage_compute, &bs, g.buf));
    HC(hipEventRecord(e1, nullptr));
    HC(hipDeviceSynchronize());
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *crpt = c_rpt;
    *ccol = c_col;
    *cval = c_val;
    *cnnz = nnz;
    *compute_s = ms * 1e-3;
    *total_s = now_s() - t0;
    return 0;
}

void brm_xcheck_free(void* p) { (void)hipFree(p); }

}  // extern "C"
