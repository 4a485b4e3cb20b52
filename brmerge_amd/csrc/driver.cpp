// brm_driver -- command-line SpGEMM driver written against
// include/brmerge_compat.hpp, i.e. against the reference's own C++
// interface (CSR, Precise_Timing, brmerge_precise, compute_flop), with the
// reference driver's protocol (BRMerge-SpGEMM/src/brmerge_precise.cpp:38-93):
// one reported warm-up call, then the mean of N timed calls, phase tables in
// the reference's layout.  The reference verifies against MKL, which is not
// available; this driver instead checks that BRMerge-Precise and
// BRMerge-Upper agree (they are bit-identical by construction).
//
//   brm_driver [--upper] [--reps N] MAT1 [MAT2]
// MAT is a MatrixMarket path or a generator spec:
//   gen:er:M:d[:seed]  gen:rmat:S[:seed]  gen:stencil:n  gen:agg:n  gen:banded:M:bw[:seed]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/brmerge_compat.hpp"

namespace {

std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    size_t b = 0;
    for (size_t e; (e = s.find(d, b)) != std::string::npos; b = e + 1) out.push_back(s.substr(b, e - b));
    out.push_back(s.substr(b));
    return out;
}

void load(const std::string& spec, CSR& A) {
    if (spec.rfind("gen:", 0) != 0) {
        A.construct(spec);
        return;
    }
    const auto f = split(spec, ':');
    auto arg = [&](size_t i, long dflt) { return f.size() > i ? atol(f[i].c_str()) : dflt; };
    int st = BRM_E_INVALID;
    int M = 0, N = 0;
    A.release();
    if (f[1] == "er") {
        M = N = (int)arg(2, 16384);
        st = brm_gen_er(M, N, (int)arg(3, 8), (uint64_t)arg(4, 0), &A.rpt, &A.col, &A.val);
    } else if (f[1] == "rmat") {
        M = N = 1 << arg(2, 16);
        st = brm_gen_rmat((int)arg(2, 16), 16, 0.57, 0.19, 0.19, (uint64_t)arg(3, 0), &A.rpt, &A.col, &A.val);
    } else if (f[1] == "stencil") {
        const int n = (int)arg(2, 32);
        M = N = n * n * n;
        st = brm_gen_stencil27(n, &A.rpt, &A.col, &A.val);
    } else if (f[1] == "agg") {
        const int n = (int)arg(2, 32), m = (n + 2) / 3;
        M = n * n * n;
        N = m * m * m;
        st = brm_gen_aggregation(n, 3, &A.rpt, &A.col, &A.val);
    } else if (f[1] == "banded") {
        M = N = (int)arg(2, 65536);
        st = brm_gen_banded(M, (int)arg(3, 64), 4, 32, (uint64_t)arg(4, 0), &A.rpt, &A.col, &A.val);
    }
    brm_compat::check(st, spec.c_str());
    A.M = M;
    A.N = N;
    A.nnz = A.rpt[M];
}

}  // namespace

int main(int argc, char** argv) {
    bool upper = false;
    int reps = 10;
    std::vector<std::string> mats;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--upper")) upper = true;
        else if (!strcmp(argv[i], "--reps") && i + 1 < argc) reps = atoi(argv[++i]);
        else mats.push_back(argv[i]);
    }
    if (mats.empty()) mats.push_back("gen:er:16384:8");
    if (mats.size() == 1) mats.push_back(mats[0]);
    try {
        CSR A, B;
        load(mats[0], A);
        if (mats[1] == mats[0]) {
            B = A;
        } else {
            load(mats[1], B);
            if (A.N < B.M) {  // trim the inner dimension as the reference driver does
                CSR t(B, A.N, B.N, 0, 0);
                B = t;
            } else if (A.N > B.M) {
                CSR t(A, A.M, B.M, 0, 0);
                A = t;
            }
        }
        printf("A %d x %d nnz %d, B %d x %d nnz %d\n", A.M, A.N, A.nnz, B.M, B.N, B.nnz);
        const long total_flop = compute_flop(A, B);
        CSR C;
        if (!upper) {
            Precise_Timing timing, bench;
            brmerge_precise(A, B, C, timing);
            timing.print(total_flop * 2.0);
            for (int i = 0; i < reps; ++i) {
                C.release();
                brmerge_precise(A, B, C, timing);
                bench += timing;
            }
            bench /= reps;
            bench.print(total_flop * 2.0);
        } else {
            Upper_Timing timing, bench;
            brmerge_upper(A, B, C, timing);
            timing.print(total_flop * 2.0);
            for (int i = 0; i < reps; ++i) {
                C.release();
                brmerge_upper(A, B, C, timing);
                bench += timing;
            }
            bench /= reps;
            bench.print(total_flop * 2.0);
        }
        CSR D;
        if (upper) {
            Precise_Timing t;
            brmerge_precise(A, B, D, t);
        } else {
            Upper_Timing t;
            brmerge_upper(A, B, D, t);
        }
        const bool same = C.nnz == D.nnz && C == D;
        printf("nnz(C) %d, Precise and Upper %s\n", C.nnz, same ? "agree: pass" : "differ: fail");
        return same ? 0 : 1;
    } catch (const std::exception& e) {
        fprintf(stderr, "brm_driver: %s\n", e.what());
        return 2;
    }
}
