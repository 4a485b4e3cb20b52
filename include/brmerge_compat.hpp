// brmerge_compat.hpp -- the reference's C++ SpGEMM interface over the
// libbrmerge.so C ABI, so that callers of BRMerge-SpGEMM compile unchanged.
//
// Replaces, for a caller, the reference's
//   inc/define.h   (mint, mdouble)                      define.h:14-15
//   inc/CSR.h + src/CSR.cpp (class CSR)                 CSR.h:7-28, CSR.cpp:102-479
//   inc/Timings.h + src/Timings.cpp                     Timings.h:4-46
//   inc/spgemm.h   brmerge_precise / brmerge_upper /    spgemm.h:220-223, 460-464,
//                  compute_flop                         468-471, 737-741, 59-66
// A caller includes this header instead of spgemm.h, CSR.h and Timings.h and
// links libbrmerge.so; the products run on the GPU.  Differences a caller
// can observe: failures throw std::runtime_error (the reference has no error
// path); CSR::operator= releases the previous arrays (the reference leaks
// them, CSR.cpp:129-141); the .mtx reader is the library's (same accepted
// formats and the same rejection of duplicate entries, CSR.cpp:259-479).
#pragma once

#include <stdio.h>
#include <string.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "brmerge.h"
#include "brmerge_gen.h"

typedef int mint;
typedef double mdouble;

namespace brm_compat {
inline void check(int st, const char* what) {
    if (st != BRM_OK) throw std::runtime_error(std::string(what) + ": " + brm_last_error());
}
}  // namespace brm_compat

class CSR {
   public:
    mint M;
    mint N;
    mint nnz;
    mint* rpt;
    mint* col;
    mdouble* val;

    CSR() : M(0), N(0), nnz(0), rpt(nullptr), col(nullptr), val(nullptr) {}
    explicit CSR(const std::string& mtx_file) : CSR() { construct(mtx_file); }
    CSR(const CSR& A) : CSR() { copy_from(A); }
    // Sub-matrix rows [M_start, M_start+M_), columns [N_start, N_start+N_) (CSR.cpp:143-179).
    CSR(const CSR& A, mint M_, mint N_, mint M_start = 0, mint N_start = 0) : CSR() {
        if (M_ + M_start > A.M || N_ + N_start > A.N) throw std::runtime_error("matrix subsect error");
        M = M_;
        N = N_;
        std::vector<mint> r(1, 0), c;
        std::vector<mdouble> v;
        for (mint i = M_start; i < M_start + M_; ++i) {
            for (mint j = A.rpt[i]; j < A.rpt[i + 1]; ++j)
                if (A.col[j] >= N_start && A.col[j] < N_start + N_) {
                    c.push_back(A.col[j] - N_start);
                    v.push_back(A.val[j]);
                }
            r.push_back((mint)c.size());
        }
        adopt(r, c, v);
    }
    // Row subset (CSR.cpp:181-212).
    CSR(const CSR& A, const std::vector<int>& rows) : CSR() {
        M = (mint)rows.size();
        N = A.N;
        std::vector<mint> r(1, 0), c;
        std::vector<mdouble> v;
        for (int i : rows) {
            if (i < 0 || i >= A.M) throw std::runtime_error("row out of range");
            c.insert(c.end(), A.col + A.rpt[i], A.col + A.rpt[i + 1]);
            v.insert(v.end(), A.val + A.rpt[i], A.val + A.rpt[i + 1]);
            r.push_back((mint)c.size());
        }
        adopt(r, c, v);
    }
    ~CSR() { release(); }

    void release() {
        delete[] rpt;
        delete[] col;
        delete[] val;
        rpt = nullptr;
        col = nullptr;
        val = nullptr;
    }

    CSR& operator=(const CSR& A) {
        if (this != &A) {
            release();
            copy_from(A);
        }
        return *this;
    }

    // Structure exact, values within 1e-9 absolute or relative; throws when
    // nnz differs (CSR.cpp:214-253).
    bool operator==(const CSR& B) {
        if (nnz != B.nnz) throw std::runtime_error("nnz not equal");
        if (M != B.M || N != B.N) return false;
        int errors = 0;
        for (mint i = 0; i <= M && errors <= 10; ++i) errors += rpt[i] != B.rpt[i];
        for (mint j = 0; j < nnz && errors <= 10; ++j) {
            const double d = std::fabs(val[j] - B.val[j]);
            errors += (col[j] != B.col[j]) || !(d < 1e-9 || d < 1e-9 * std::fabs(val[j]));
        }
        return errors == 0;
    }

    void construct(const std::string& mtx_file) {
        release();
        int m = 0, n = 0;
        int64_t z = 0;
        brm_compat::check(brm_read_mtx(mtx_file.c_str(), &m, &n, &z, &rpt, &col, &val), "brm_read_mtx");
        M = m;
        N = n;
        nnz = (mint)z;
    }

   private:
    void copy_from(const CSR& A) {
        M = A.M;
        N = A.N;
        nnz = A.nnz;
        rpt = new mint[M + 1];
        col = new mint[nnz > 0 ? nnz : 1];
        val = new mdouble[nnz > 0 ? nnz : 1];
        memcpy(rpt, A.rpt, sizeof(mint) * (M + 1));
        if (nnz > 0) {
            memcpy(col, A.col, sizeof(mint) * nnz);
            memcpy(val, A.val, sizeof(mdouble) * nnz);
        }
    }
    void adopt(const std::vector<mint>& r, const std::vector<mint>& c, const std::vector<mdouble>& v) {
        nnz = (mint)c.size();
        rpt = new mint[r.size()];
        col = new mint[nnz > 0 ? nnz : 1];
        val = new mdouble[nnz > 0 ? nnz : 1];
        memcpy(rpt, r.data(), sizeof(mint) * r.size());
        if (nnz > 0) {
            memcpy(col, c.data(), sizeof(mint) * nnz);
            memcpy(val, v.data(), sizeof(mdouble) * nnz);
        }
    }
};

namespace brm_compat {
// Prints a phase table in the reference's layout (ms, % of total, then
// GFLOPS per phase) for any list of (name, seconds) pairs.
inline void print_phases(const char* const* names, const double* secs, int n, double total, int sum_upto,
                         double total_flop) {
    const double g = total_flop / 1e9;
    double sum = 0;
    for (int i = 0; i < sum_upto; ++i) sum += secs[i];
    printf("total flop %lf\n", total_flop);
    printf("time(ms):\n");
    for (int i = 0; i < n; ++i) printf("    %-20s%8.3lfms %6.2lf%%\n", names[i], 1000 * secs[i], secs[i] / total * 100);
    printf("    %-20s%8.3lfms %6.2lf%%\n", "sum_total", 1000 * sum, sum / total * 100);
    printf("perf(GFLOPS):\n");
    for (int i = 0; i < n; ++i) printf("    %-20s%6.2lf\n", names[i], g / secs[i]);
    printf("    %-20s%6.2lf\n", "total", g / total);
}
}  // namespace brm_compat

class Precise_Timing {
   public:
    bool measure_separate = true;
    bool measure_whole = true;
    double pre_allocate = 0, compute_flop = 0, prefix_sum_flop = 0, load_balance = 0, symbolic = 0,
           prefix_sum_nnz = 0, allocate_C = 0, compute = 0, cleanup = 0, total = 0;
    double* fields() { return &pre_allocate; }
    const double* fields() const { return &pre_allocate; }
    static constexpr int kFields = 10;
    void operator+=(const Precise_Timing& a) {
        for (int i = 0; i < kFields; ++i) fields()[i] += a.fields()[i];
    }
    void operator/=(const double n) {
        for (int i = 0; i < kFields; ++i) fields()[i] /= n;
    }
    void print(const double total_flop) {
        static const char* const names[] = {"pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance",
                                            "symbolic", "prefix_sum_nnz", "allocate_C", "compute", "cleanup"};
        if (measure_separate) brm_compat::print_phases(names, fields(), 9, total, 8, total_flop);
    }
    void reg_print(const double total_flop) { printf("%lf\n", total_flop / 1e9 / total); }
    void load(const brm_precise_timing& t) {
        pre_allocate = t.pre_allocate + t.h2d;
        compute_flop = t.compute_flop;
        prefix_sum_flop = t.prefix_sum_flop;
        load_balance = t.load_balance;
        symbolic = t.symbolic;
        prefix_sum_nnz = t.prefix_sum_nnz;
        allocate_C = t.allocate_C;
        compute = t.compute;
        cleanup = t.cleanup + t.d2h;
        total = t.total;
    }
};

class Upper_Timing {
   public:
    bool measure_separate = true;
    bool measure_whole = true;
    double thread = 0, pre_allocate = 0, compute_flop = 0, prefix_sum_flop = 0, load_balance = 0, allocate_hook = 0,
           compute = 0, prefix_sum_nnz = 0, allocate_C = 0, copy = 0, cleanup = 0, total = 0;
    double* fields() { return &thread; }
    const double* fields() const { return &thread; }
    static constexpr int kFields = 12;
    void operator+=(const Upper_Timing& a) {
        for (int i = 0; i < kFields; ++i) fields()[i] += a.fields()[i];
    }
    void operator/=(const double n) {
        for (int i = 0; i < kFields; ++i) fields()[i] /= n;
    }
    void print(const double total_flop) {
        static const char* const names[] = {"thread", "pre_allocate", "compute_flop", "prefix_sum_flop",
                                            "load_balance", "allocate_hook", "compute", "prefix_sum_nnz",
                                            "allocate_C", "copy", "cleanup"};
        if (measure_separate) brm_compat::print_phases(names, fields(), 11, total, 11, total_flop);
    }
    void reg_print(const double total_flop) { printf("%le\n", total_flop / 1e9 / total); }
    void load(const brm_upper_timing& t) {
        thread = t.thread;
        pre_allocate = t.pre_allocate + t.h2d;
        compute_flop = t.compute_flop;
        prefix_sum_flop = t.prefix_sum_flop;
        load_balance = t.load_balance;
        allocate_hook = t.allocate_hook;
        compute = t.compute;
        prefix_sum_nnz = t.prefix_sum_nnz;
        allocate_C = t.allocate_C;
        copy = t.copy;
        cleanup = t.cleanup + t.d2h;
        total = t.total;
    }
};

// spgemm.h:468-471 / 220-223: raw-pointer forms.  Outputs are new[] arrays
// owned by the caller, as in the reference.
inline void brmerge_precise(mint* arpt, mint* acol, mdouble* aval, mint* brpt, mint* bcol, mdouble* bval,
                            mint** crpt_, mint** ccol_, mdouble** cval_, mint M, mint K, mint N, mint* cnnz_,
                            Precise_Timing& timing) {
    brm_precise_timing t;
    brm_compat::check(brm_spgemm_precise_i32(arpt, acol, aval, brpt, bcol, bval, crpt_, ccol_, cval_, M, K, N,
                                             cnnz_, &t),
                      "brmerge_precise");
    timing.load(t);
}

inline void brmerge_upper(mint* arpt, mint* acol, mdouble* aval, mint* brpt, mint* bcol, mdouble* bval,
                          mint** crpt_, mint** ccol_, mdouble** cval_, mint M, mint K, mint N, mint* cnnz_,
                          Upper_Timing& timing) {
    brm_upper_timing t;
    brm_compat::check(brm_spgemm_upper_i32(arpt, acol, aval, brpt, bcol, bval, crpt_, ccol_, cval_, M, K, N,
                                           cnnz_, &t),
                      "brmerge_upper");
    timing.load(t);
}

// spgemm.h:737-741 / 460-464: CSR forms (C.M = A.M, C.N = B.N).
inline void brmerge_precise(const CSR& A, const CSR& B, CSR& C, Precise_Timing& timing) {
    C.M = A.M;
    C.N = B.N;
    brmerge_precise(A.rpt, A.col, A.val, B.rpt, B.col, B.val, &C.rpt, &C.col, &C.val, A.M, A.N, B.N, &C.nnz, timing);
}

inline void brmerge_upper(const CSR& A, const CSR& B, CSR& C, Upper_Timing& timing) {
    C.M = A.M;
    C.N = B.N;
    brmerge_upper(A.rpt, A.col, A.val, B.rpt, B.col, B.val, &C.rpt, &C.col, &C.val, A.M, A.N, B.N, &C.nnz, timing);
}

// spgemm.h:59-66: number of intermediate products of A*B.
inline long compute_flop(const CSR& A, const CSR& B) {
    int64_t total = 0;
    brm_compat::check(brm_flop_count(A.rpt, A.col, B.rpt, A.M, B.M, &total), "compute_flop");
    return (long)total;
}
