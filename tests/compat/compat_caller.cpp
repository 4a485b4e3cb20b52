// A caller written against the reference's own C++ SpGEMM interface
// (CSR, Precise_Timing, Upper_Timing, brmerge_precise / brmerge_upper in
// both overloads, compute_flop), compiled against include/brmerge_compat.hpp
// instead of BRMerge-SpGEMM/inc/spgemm.h.  Used by tests/test_gpu_compat.py.
//
//   compat_caller A.mtx B.mtx OUT_PREFIX
// Reads A and B (CSR(path), CSR.cpp:259-479), prints compute_flop, runs
//   1. brmerge_precise(const CSR&, const CSR&, CSR&, Precise_Timing&)   spgemm.h:737-741
//   2. brmerge_upper(const CSR&, const CSR&, CSR&, Upper_Timing&)       spgemm.h:460-464
//   3. the raw-pointer brmerge_precise                                  spgemm.h:468-471
//   4. the raw-pointer brmerge_upper                                    spgemm.h:220-223
// prints both timing tables and reg_print lines, checks CSR::operator==
// semantics (CSR.cpp:214-253) and release/copy/assignment, and writes the C
// of form 1 as OUT_PREFIX.{rpt,col,val} (raw int32 / int32 / float64).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>

#include "brmerge_compat.hpp"

static int fails = 0;
#define EXPECT(cond)                                                  \
    do {                                                              \
        if (!(cond)) {                                                \
            fprintf(stderr, "EXPECT failed line %d: %s\n", __LINE__, #cond); \
            ++fails;                                                  \
        }                                                             \
    } while (0)

static bool same_bits(const CSR& X, const CSR& Y) {
    if (X.M != Y.M || X.N != Y.N || X.nnz != Y.nnz) return false;
    return !memcmp(X.rpt, Y.rpt, sizeof(mint) * (X.M + 1)) && !memcmp(X.col, Y.col, sizeof(mint) * X.nnz) &&
           !memcmp(X.val, Y.val, sizeof(mdouble) * X.nnz);
}

static void dump(const std::string& path, const void* p, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: compat_caller A.mtx B.mtx OUT_PREFIX\n");
        return 2;
    }
    try {
        CSR A(argv[1]), B(argv[2]);
        const long flop = compute_flop(A, B);
        printf("flop %ld\n", flop);

        CSR C1, C2;
        Precise_Timing tp;
        brmerge_precise(A, B, C1, tp);
        EXPECT(C1.M == A.M && C1.N == B.N);
        printf("== precise\n");
        tp.print(2.0 * flop);
        tp.reg_print(2.0 * flop);
        Upper_Timing tu;
        brmerge_upper(A, B, C2, tu);
        printf("== upper\n");
        tu.print(2.0 * flop);
        tu.reg_print(2.0 * flop);
        EXPECT(same_bits(C1, C2));

        // raw-pointer overloads: caller-owned new[] outputs
        mint *crpt = nullptr, *ccol = nullptr, cnnz = -1;
        mdouble* cval = nullptr;
        Precise_Timing t3;
        brmerge_precise(A.rpt, A.col, A.val, B.rpt, B.col, B.val, &crpt, &ccol, &cval, A.M, A.N, B.N, &cnnz, t3);
        EXPECT(cnnz == C1.nnz);
        EXPECT(!memcmp(crpt, C1.rpt, sizeof(mint) * (A.M + 1)));
        EXPECT(cnnz == 0 || (!memcmp(ccol, C1.col, sizeof(mint) * cnnz) && !memcmp(cval, C1.val, sizeof(mdouble) * cnnz)));
        delete[] crpt;
        delete[] ccol;
        delete[] cval;
        Upper_Timing t4;
        brmerge_upper(A.rpt, A.col, A.val, B.rpt, B.col, B.val, &crpt, &ccol, &cval, A.M, A.N, B.N, &cnnz, t4);
        EXPECT(cnnz == C1.nnz);
        EXPECT(cnnz == 0 || !memcmp(cval, C1.val, sizeof(mdouble) * cnnz));
        delete[] crpt;
        delete[] ccol;
        delete[] cval;

        // operator== : structure exact, values within 1e-9; nnz mismatch throws
        CSR D(C1);
        EXPECT(D == C1);
        if (D.nnz > 0) {
            D.val[D.nnz / 2] += 1e-3 * (1.0 + (D.val[D.nnz / 2] < 0 ? -D.val[D.nnz / 2] : D.val[D.nnz / 2]));
            EXPECT(!(D == C1));
        }
        CSR E(C1, C1.M > 1 ? C1.M - 1 : C1.M, C1.N);  // sub-matrix constructor
        bool threw = false;
        try {
            (void)(E == C1);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        EXPECT(threw == (E.nnz != C1.nnz));
        // assignment and release, then reuse of the released CSR as an output
        D = C1;
        EXPECT(same_bits(D, C1));
        D.release();
        EXPECT(D.rpt == nullptr && D.col == nullptr && D.val == nullptr);
        brmerge_precise(A, B, D, tp);
        EXPECT(same_bits(D, C1));

        dump(std::string(argv[3]) + ".rpt", C1.rpt, sizeof(mint) * (C1.M + 1));
        dump(std::string(argv[3]) + ".col", C1.col, sizeof(mint) * (C1.nnz > 0 ? C1.nnz : 0));
        dump(std::string(argv[3]) + ".val", C1.val, sizeof(mdouble) * (C1.nnz > 0 ? C1.nnz : 0));
        printf("nnz %d M %d N %d checks_failed %d\n", C1.nnz, C1.M, C1.N, fails);
    } catch (const std::exception& e) {
        fprintf(stderr, "compat_caller: %s\n", e.what());
        return 2;
    }
    return fails ? 1 : 0;
}
