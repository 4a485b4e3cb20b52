/*
 * brmerge_gen.h -- deterministic synthetic CSR generators, MatrixMarket
 * input and a host CSR transpose (part of libbrmerge.so, host code only).
 *
 * Every generator draws from counter-based streams (SplitMix64 seeding
 * xoshiro256**, one stream per row / edge / entry), so the output depends
 * only on the arguments and never on the thread count.  Outputs are
 * allocated with new[] and owned by the caller (brm_host_free_i32 /
 * brm_host_free_f64).  Columns of every row are strictly increasing.
 * Values drawn "U[-1,1]" are 2u-1 with u = (x >> 11) * 2^-53.
 */
#ifndef BRMERGE_GEN_H_
#define BRMERGE_GEN_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Erdos-Renyi: each row holds d distinct uniform columns of [0, N) (Floyd's
 * sampling), sorted, values U[-1,1]. */
int brm_gen_er(int M, int N, int d, uint64_t seed, int **rpt, int **col, double **val);

/* As brm_gen_er with a per-row count drawn uniformly from [dmin, dmax]. */
int brm_gen_er_var(int M, int N, int dmin, int dmax, uint64_t seed, int **rpt, int **col, double **val);

/* R-MAT: 2^scale x 2^scale, edge_factor * 2^scale edges, quadrant
 * probabilities a, b, c (d = 1-a-b-c) at every level, no noise; duplicate
 * edges removed, self-loops kept; values U[-1,1] in (row, col) order. */
int brm_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 int **rpt, int **col, double **val);

/* 27-point stencil on an n x n x n grid (row = (z*n + y)*n + x), 26 on the
 * diagonal and -1 on the other points. */
int brm_gen_stencil27(int n, int **rpt, int **col, double **val);

/* Aggregation prolongator of the same grid: one 1.0 per row in column
 * ((z/g)*m + y/g)*m + x/g with m = ceil(n/g).  Has m^3 columns. */
int brm_gen_aggregation(int n, int g, int **rpt, int **col, double **val);

/* Banded: row i holds d distinct columns of [i-bw, i+bw] clipped to [0, M),
 * d uniform in [dmin, dmax] (clipped to the band width), values U[-1,1]. */
int brm_gen_banded(int M, int bw, int dmin, int dmax, uint64_t seed, int **rpt, int **col, double **val);

/* Transpose of an M x N CSR matrix (rows of the result list the source rows
 * in ascending order). */
int brm_csr_transpose(int M, int N, const int *rpt, const int *col, const double *val,
                      int **trpt, int **tcol, double **tval);

/* MatrixMarket coordinate reader with the semantics of the reference's
 * CSR::construct (BRMerge-SpGEMM/src/CSR.cpp:259-479): general / symmetric /
 * skew-symmetric (hermitian rejected); real, integer, pattern (values 1.0)
 * and complex (real part kept); 1-based indices; entries sorted by
 * (row, col); duplicate entries rejected (BRM_E_INVALID). */
int brm_read_mtx(const char *path, int *M, int *N, int64_t *nnz, int **rpt, int **col, double **val);

/* Writes a general real coordinate MatrixMarket file (17 significant digits). */
int brm_write_mtx(const char *path, int M, int N, const int *rpt, const int *col, const double *val);

#ifdef __cplusplus
}
#endif
#endif /* BRMERGE_GEN_H_ */
