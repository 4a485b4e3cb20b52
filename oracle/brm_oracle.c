/*
 * brm_oracle.c -- CPU restatement of BRMerge SpGEMM (Precise and Upper).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library (brmerge_amd/), which must fail loudly when
 * its HIP extension is missing instead of falling back here.
 *
 * It restates, in plain C + OpenMP, the algorithm of
 *   BRMerge-SpGEMM/inc/spgemm.h   (compute_flop :31-56, upper_log :69-80,
 *                                  seq/para/opt_prefix_sum :138-187,
 *                                  binary_search_approx :189-215,
 *                                  brmerge_upper :220-457,
 *                                  brmerge_precise :468-735)
 * with three deliberate, result-neutral deviations:
 *   - row pointers of C and all flop accumulators are int64 (the reference's
 *     `int thread_flop` / `int cnnz` overflow on R-MAT scale 20, spgemm.h:12,35,417,571);
 *   - the symbolic hash key product is unsigned (spgemm.h:540 is signed and
 *     undefined above ~2.0e7 columns; identical for smaller keys);
 *   - buffers are allocated with malloc and freed with brmo_free.
 * The merge order (pairs (2j,2j+1) per level, odd list carried, equal
 * columns summed as src1+src2, empty lists keeping their positions) is the
 * reference's exactly, so results are bit-identical to it.
 *
 * Parity status: pinned by tests/test_oracle.py against hand-derived
 * known answers for the merge bracketing (tests/golden/kat_bracketing.json),
 * a recursive tree-definition brute force, scipy's pattern product and FP64
 * product, and closed forms (A = I, one entry per row, flop totals).
 * The reference itself cannot be compiled here (its utils.h:3 needs TBB's
 * scalable_allocator.h, absent from this image).
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define BRMO_HASH_SCALE 107u   /* define.h:12 */
#define BRMO_SYMBOLIC_SCALE 1  /* define.h:18 */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void brmo_free(void *p) { free(p); }

int brmo_max_threads(void) { return omp_get_max_threads(); }

/* spgemm.h:69-80: next power of two >= num; upper_log(0) == 2, upper_log(1) == 1. */
int brmo_upper_log(int num) {
    int a = num, base = 0;
    while ((a = a >> 1)) base++;
    if (num != (1 << base)) return 1 << (base + 1);
    return 1 << base;
}

/* spgemm.h:31-56 with int64 accumulation.  nnzrA[i] = row length of A,
 * floprC[i] = number of intermediate products of row i. */
int64_t brmo_compute_flop(const int *arpt, const int *acol, const int *brpt, int M,
                          int *nnzrA, int64_t *floprC, int nthreads) {
    int64_t total = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        int64_t tf = 0;
#pragma omp for
        for (int i = 0; i < M; i++) {
            int64_t s = 0;
            if (nnzrA) nnzrA[i] = arpt[i + 1] - arpt[i];
            for (int j = arpt[i]; j < arpt[i + 1]; j++) s += brpt[acol[j] + 1] - brpt[acol[j]];
            if (floprC) floprC[i] = s;
            tf += s;
        }
#pragma omp critical
        total += tf;
    }
    return total;
}

/* spgemm.h:138-145 / 147-177 / 179-187: exclusive prefix sum of N values into
 * out[0..N].  Sequential below 8192 entries, two-pass blocked above. */
static void prefix_sum_i64(const int64_t *in, int64_t *out, int N, int tnum) {
    if (N < (1 << 13) || tnum <= 1) {
        out[0] = 0;
        for (int i = 0; i < N; i++) out[i + 1] = out[i] + in[i];
        return;
    }
    int each_n = N / tnum;
    int64_t *partial = (int64_t *)malloc(sizeof(int64_t) * tnum);
#pragma omp parallel num_threads(tnum)
    {
        int tid = omp_get_thread_num();
        int nt = omp_get_num_threads();
        if (nt == tnum) {
            int start = each_n * tid;
            int end = (tid < tnum - 1) ? start + each_n : N;
            out[start] = 0;
            for (int i = start; i < end - 1; ++i) out[i + 1] = out[i] + in[i];
            partial[tid] = out[end - 1] + in[end - 1];
#pragma omp barrier
            int64_t offset = 0;
            for (int i = 0; i < tid; ++i) offset += partial[i];
            for (int i = start; i < end; ++i) out[i] += offset;
            if (tid == tnum - 1) out[N] = out[N - 1] + in[N - 1];
        }
    }
    free(partial);
}

void brmo_prefix_sum(const int64_t *in, int64_t *out, int N, int tnum) { prefix_sum_i64(in, out, N, tnum); }

/* spgemm.h:189-215: index m with arr[m] <= elem < arr[m+1] (approximately). */
int brmo_binary_search_approx(const int64_t *arr, int N, int64_t elem) {
    int s = 0, e = N - 1;
    if (elem >= arr[N - 1]) return N - 1;
    while (1) {
        int m = (s + e) / 2;
        if (arr[m] == elem) return m;
        else if (arr[m] < elem) {
            if (elem < arr[m + 1]) return m;
            s = m;
        } else {
            if (arr[m - 1] < elem) return m - 1;
            e = m;
        }
    }
}

/* spgemm.h:247-256 / 494-503: flop-balanced contiguous row ranges. */
static void load_balance(const int64_t *scan_flop, int M, int64_t total, int T, int64_t *sec) {
    float divide_flops = (float)total / (float)T;
    sec[0] = 0;
    for (int i = 1; i < T; i++) sec[i] = brmo_binary_search_approx(scan_flop, M + 1, (int64_t)(divide_flops * i));
    sec[T] = M;
    for (int i = 1; i <= T; i++) if (sec[i] < sec[i - 1]) sec[i] = sec[i - 1];
}

/*
 * One row: EXPAND (spgemm.h:618-624) then MERGE (spgemm.h:639-707).
 * Returns the merged length; the result is left in (*res_col, *res_val),
 * which point into ping or pong.
 */
static int64_t merge_row(const int *arpt, const int *acol, const double *aval,
                         const int *brpt, const int *bcol, const double *bval, int row,
                         int *col_ping, double *val_ping, int64_t *off_ping,
                         int *col_pong, double *val_pong, int64_t *off_pong,
                         int **res_col, double **res_val) {
    int *dst_col = col_ping, *src_col;
    double *dst_val = val_ping, *src_val;
    int64_t *dst_off = off_ping, *src_off;
    int64_t left_num_list = 0, buf = 0;
    dst_off[0] = 0;
    for (int a = arpt[row]; a < arpt[row + 1]; a++) {
        double av = aval[a];
        for (int b = brpt[acol[a]]; b < brpt[acol[a] + 1]; b++) {
            dst_col[buf] = bcol[b];
            dst_val[buf++] = av * bval[b];
        }
        dst_off[++left_num_list] = buf;
    }
    if (left_num_list == 0) { *res_col = col_ping; *res_val = val_ping; return 0; }

    src_col = col_ping; src_val = val_ping; src_off = off_ping;
    dst_col = col_pong; dst_val = val_pong; dst_off = off_pong;
    while (left_num_list != 1) {
        int64_t inner = left_num_list;
        int64_t src_li = 0, dst_li = 0;
        left_num_list = 0;
        dst_off[0] = 0;
        while (inner) {
            int64_t d = dst_off[dst_li];
            if (inner >= 2) {
                int64_t s1 = src_off[src_li], e1 = src_off[src_li + 1];
                int64_t s2 = src_off[src_li + 1], e2 = src_off[src_li + 2];
                while (s1 < e1 && s2 < e2) {
                    if (src_col[s1] == src_col[s2]) {
                        dst_col[d] = src_col[s1];
                        dst_val[d++] = src_val[s1++] + src_val[s2++];
                    } else if (src_col[s1] < src_col[s2]) {
                        dst_col[d] = src_col[s1];
                        dst_val[d++] = src_val[s1++];
                    } else {
                        dst_col[d] = src_col[s2];
                        dst_val[d++] = src_val[s2++];
                    }
                }
                if (s1 < e1) {
                    memcpy(dst_col + d, src_col + s1, (size_t)(e1 - s1) * sizeof(int));
                    memcpy(dst_val + d, src_val + s1, (size_t)(e1 - s1) * sizeof(double));
                    d += e1 - s1;
                } else if (s2 < e2) {
                    memcpy(dst_col + d, src_col + s2, (size_t)(e2 - s2) * sizeof(int));
                    memcpy(dst_val + d, src_val + s2, (size_t)(e2 - s2) * sizeof(double));
                    d += e2 - s2;
                }
                src_li += 2;
                left_num_list++;
                inner -= 2;
                dst_off[++dst_li] = d;
            } else { /* inner == 1: carry the odd list */
                int64_t s1 = src_off[src_li], n1 = src_off[src_li + 1] - s1;
                memcpy(dst_col + d, src_col + s1, (size_t)n1 * sizeof(int));
                memcpy(dst_val + d, src_val + s1, (size_t)n1 * sizeof(double));
                d += n1;
                src_li++;
                left_num_list++;
                inner--;
                dst_off[++dst_li] = d;
            }
        }
        { int *t = src_col; src_col = dst_col; dst_col = t; }
        { double *t = src_val; src_val = dst_val; dst_val = t; }
        { int64_t *t = src_off; src_off = dst_off; dst_off = t; }
    }
    *res_col = src_col;
    *res_val = src_val;
    return src_off[1];
}

/* Phase times in seconds, in the order of Precise_Timing (Timings.h:27-46). */
enum { PT_PRE_ALLOCATE, PT_COMPUTE_FLOP, PT_PREFIX_SUM_FLOP, PT_LOAD_BALANCE, PT_SYMBOLIC,
       PT_PREFIX_SUM_NNZ, PT_ALLOCATE_C, PT_COMPUTE, PT_CLEANUP, PT_TOTAL, PT_N };

/*
 * spgemm.h:468-735.  Outputs are malloc'd (free with brmo_free); crpt is
 * int64[M+1].  Returns 0, or -1 on allocation failure.
 */
int brmo_brmerge_precise(const int *arpt, const int *acol, const double *aval,
                         const int *brpt, const int *bcol, const double *bval,
                         int M, int K, int N, int64_t **crpt_, int **ccol_, double **cval_,
                         int64_t *cnnz_, int nthreads, double *tm) {
    (void)K; (void)N;
    double t0, t1, tt[PT_N] = {0};
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int T = omp_get_max_threads();
    t0 = t1 = now_s();
    int *nnzrA = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    int64_t *floprC = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M > 0 ? M : 1));
    int64_t *scan_floprC = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *crpt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *sec = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
    int64_t *max_flop = (int64_t *)malloc(sizeof(int64_t) * T);
    int *max_nnzrA = (int *)malloc(sizeof(int) * T);
    if (!nnzrA || !floprC || !scan_floprC || !crpt || !sec || !max_flop || !max_nnzrA) return -1;
    tt[PT_PRE_ALLOCATE] = now_s() - t0;

    t0 = now_s();
    int64_t total = brmo_compute_flop(arpt, acol, brpt, M, nnzrA, floprC, 0);
    tt[PT_COMPUTE_FLOP] = now_s() - t0;

    t0 = now_s();
    prefix_sum_i64(floprC, scan_floprC, M, T);
    tt[PT_PREFIX_SUM_FLOP] = now_s() - t0;

    t0 = now_s();
    load_balance(scan_floprC, M, total, T, sec);
    tt[PT_LOAD_BALANCE] = now_s() - t0;

    /* SYMBOLIC, spgemm.h:505-562: linear-probing hash per row. */
    t0 = now_s();
    int fail = 0;
#pragma omp parallel num_threads(T)
    {
        int rid = omp_get_thread_num();
        int rs = (int)sec[rid], re = (int)sec[rid + 1];
        int tmax_n = 0;
        int64_t tmax_f = 0;
        for (int i = rs; i < re; i++) {
            if (tmax_n < nnzrA[i]) tmax_n = nnzrA[i];
            if (tmax_f < floprC[i]) tmax_f = floprC[i];
        }
        max_nnzrA[rid] = tmax_n;
        max_flop[rid] = tmax_f;
        int cap = brmo_upper_log((int)(tmax_f > 0x40000000 ? 0x40000000 : tmax_f));
        int *ht = (int *)malloc(sizeof(int) * (size_t)cap);
        if (!ht) {
#pragma omp atomic write
            fail = 1;
        } else {
            for (int row = rs; row < re; row++) {
                int ht_size = brmo_upper_log((int)floprC[row] * BRMO_SYMBOLIC_SCALE);
                for (int j = 0; j < ht_size; j++) ht[j] = -1;
                int64_t nnz = 0;
                for (int j = arpt[row]; j < arpt[row + 1]; j++) {
                    int ac = acol[j];
                    for (int k = brpt[ac]; k < brpt[ac + 1]; k++) {
                        int key = bcol[k];
                        unsigned h = ((unsigned)key * BRMO_HASH_SCALE) & (unsigned)(ht_size - 1);
                        for (;;) {
                            if (ht[h] == key) break;
                            if (ht[h] == -1) { ht[h] = key; nnz++; break; }
                            h = (h + 1) & (unsigned)(ht_size - 1);
                        }
                    }
                }
                floprC[row] = nnz;
            }
            free(ht);
        }
    }
    if (fail) return -1;
    tt[PT_SYMBOLIC] = now_s() - t0;

    t0 = now_s();
    prefix_sum_i64(floprC, crpt, M, T);
    tt[PT_PREFIX_SUM_NNZ] = now_s() - t0;

    t0 = now_s();
    int64_t cnnz = crpt[M];
    int *ccol = (int *)malloc(sizeof(int) * (size_t)(cnnz > 0 ? cnnz : 1));
    double *cval = (double *)malloc(sizeof(double) * (size_t)(cnnz > 0 ? cnnz : 1));
    if (!ccol || !cval) return -1;
    tt[PT_ALLOCATE_C] = now_s() - t0;

    /* NUMERIC, spgemm.h:580-720. */
    t0 = now_s();
#pragma omp parallel num_threads(T)
    {
        int rid = omp_get_thread_num();
        int rs = (int)sec[rid], re = (int)sec[rid + 1];
        size_t mf = (size_t)(max_flop[rid] > 0 ? max_flop[rid] : 1);
        size_t mn = (size_t)max_nnzrA[rid] + 1;
        int *cpi = (int *)malloc(sizeof(int) * mf), *cpo = (int *)malloc(sizeof(int) * mf);
        double *vpi = (double *)malloc(sizeof(double) * mf), *vpo = (double *)malloc(sizeof(double) * mf);
        int64_t *opi = (int64_t *)malloc(sizeof(int64_t) * mn), *opo = (int64_t *)malloc(sizeof(int64_t) * mn);
        if (!cpi || !cpo || !vpi || !vpo || !opi || !opo) {
#pragma omp atomic write
            fail = 1;
        } else {
            for (int row = rs; row < re; row++) {
                int *rc; double *rv;
                int64_t n = merge_row(arpt, acol, aval, brpt, bcol, bval, row, cpi, vpi, opi, cpo, vpo, opo, &rc, &rv);
                if (n != floprC[row]) {
#pragma omp atomic write
                    fail = 2;
                }
                if (n > 0) {
                    memcpy(ccol + crpt[row], rc, (size_t)n * sizeof(int));
                    memcpy(cval + crpt[row], rv, (size_t)n * sizeof(double));
                }
            }
        }
        free(cpi); free(cpo); free(vpi); free(vpo); free(opi); free(opo);
    }
    tt[PT_COMPUTE] = now_s() - t0;

    t0 = now_s();
    free(nnzrA); free(floprC); free(scan_floprC); free(sec); free(max_flop); free(max_nnzrA);
    tt[PT_CLEANUP] = now_s() - t0;
    tt[PT_TOTAL] = now_s() - t1;
    if (fail) { free(ccol); free(cval); free(crpt); return -fail; }
    *crpt_ = crpt; *ccol_ = ccol; *cval_ = cval; *cnnz_ = cnnz;
    if (tm) memcpy(tm, tt, sizeof(tt));
    return 0;
}

/*
 * spgemm.h:220-457.  Numeric into per-thread C-bar sized by the flop upper
 * bound, then prefix, allocate and compact.  Same output contract as
 * brmo_brmerge_precise.
 */
int brmo_brmerge_upper(const int *arpt, const int *acol, const double *aval,
                       const int *brpt, const int *bcol, const double *bval,
                       int M, int K, int N, int64_t **crpt_, int **ccol_, double **cval_,
                       int64_t *cnnz_, int nthreads, double *tm) {
    (void)K; (void)N;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int T = omp_get_max_threads();
    double t0 = now_s();
    int *nnzrA = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    int64_t *floprC = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M > 0 ? M : 1));
    int64_t *scan_floprC = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *crpt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *sec = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
    int **colbar = (int **)calloc((size_t)T, sizeof(int *));
    double **valbar = (double **)calloc((size_t)T, sizeof(double *));
    if (!nnzrA || !floprC || !scan_floprC || !crpt || !sec || !colbar || !valbar) return -1;
    int64_t total = brmo_compute_flop(arpt, acol, brpt, M, nnzrA, floprC, 0);
    prefix_sum_i64(floprC, scan_floprC, M, T);
    load_balance(scan_floprC, M, total, T, sec);
    int fail = 0;
    double tc = now_s();
#pragma omp parallel num_threads(T)
    {
        int rid = omp_get_thread_num();
        int rs = (int)sec[rid], re = (int)sec[rid + 1];
        int64_t base = scan_floprC[rs], local = scan_floprC[re] - base;
        int mn = 0;
        int64_t mf = 0;
        for (int i = rs; i < re; i++) {
            if (mn < nnzrA[i]) mn = nnzrA[i];
            if (mf < floprC[i]) mf = floprC[i];
        }
        size_t mfs = (size_t)(mf > 0 ? mf : 1), mns = (size_t)mn + 1;
        int *cpi = (int *)malloc(sizeof(int) * mfs), *cpo = (int *)malloc(sizeof(int) * mfs);
        double *vpi = (double *)malloc(sizeof(double) * mfs), *vpo = (double *)malloc(sizeof(double) * mfs);
        int64_t *opi = (int64_t *)malloc(sizeof(int64_t) * mns), *opo = (int64_t *)malloc(sizeof(int64_t) * mns);
        colbar[rid] = (int *)malloc(sizeof(int) * (size_t)(local > 0 ? local : 1));
        valbar[rid] = (double *)malloc(sizeof(double) * (size_t)(local > 0 ? local : 1));
        if (!cpi || !cpo || !vpi || !vpo || !opi || !opo || !colbar[rid] || !valbar[rid]) {
#pragma omp atomic write
            fail = 1;
        } else {
            for (int row = rs; row < re; row++) {
                int *rc; double *rv;
                int64_t n = merge_row(arpt, acol, aval, brpt, bcol, bval, row, cpi, vpi, opi, cpo, vpo, opo, &rc, &rv);
                floprC[row] = n; /* spgemm.h:398 */
                if (n > 0) {
                    memcpy(colbar[rid] + scan_floprC[row] - base, rc, (size_t)n * sizeof(int));
                    memcpy(valbar[rid] + scan_floprC[row] - base, rv, (size_t)n * sizeof(double));
                }
            }
        }
        free(cpi); free(cpo); free(vpi); free(vpo); free(opi); free(opo);
    }
    double tcomp = now_s() - tc;
    int64_t cnnz = 0;
    int *ccol = NULL;
    double *cval = NULL;
    if (!fail) {
        prefix_sum_i64(floprC, crpt, M, T); /* spgemm.h:413 */
        cnnz = crpt[M];
        ccol = (int *)malloc(sizeof(int) * (size_t)(cnnz > 0 ? cnnz : 1));
        cval = (double *)malloc(sizeof(double) * (size_t)(cnnz > 0 ? cnnz : 1));
        if (!ccol || !cval) fail = 1;
    }
    if (!fail) {
#pragma omp parallel num_threads(T)
        {   /* COMPACT, spgemm.h:427-445 */
            int rid = omp_get_thread_num();
            int rs = (int)sec[rid], re = (int)sec[rid + 1];
            int64_t base = scan_floprC[rs];
            for (int row = rs; row < re; row++) {
                int64_t n = floprC[row];
                if (n > 0) {
                    memcpy(ccol + crpt[row], colbar[rid] + scan_floprC[row] - base, (size_t)n * sizeof(int));
                    memcpy(cval + crpt[row], valbar[rid] + scan_floprC[row] - base, (size_t)n * sizeof(double));
                }
            }
        }
    }
    for (int t = 0; t < T; t++) { free(colbar[t]); free(valbar[t]); }
    free(colbar); free(valbar); free(nnzrA); free(floprC); free(scan_floprC); free(sec);
    if (fail) { free(ccol); free(cval); free(crpt); return -1; }
    *crpt_ = crpt; *ccol_ = ccol; *cval_ = cval; *cnnz_ = cnnz;
    if (tm) { tm[0] = tcomp; tm[1] = now_s() - t0; }
    return 0;
}

/*
 * Existing-SpGEMM/inc/hash_mult.h:558-608 (hash_numeric, sortOutput) and
 * :519-552 (sort_and_store_table2mat): Hash SpGEMM's numeric phase, the
 * paper's comparison baseline.  Per row, a linear-probing table keyed by
 * column, hash (key * 107) & (size - 1) (HASH_SCAL, :23), of
 * MIN_HT_N << (bin_id - 1) slots (BIN.h:157-177, MIN_HT_N 8 outside KNL
 * builds, :34-35; bin_id from the row's products, capped at the column
 * count); products visited in A-entry order, then B-entry order, each
 * rounded (multiplies) and added onto its column's value as
 * t_val + ht_value (:588); the occupied slots sorted by column.  The value
 * of each column is therefore the sequential left fold of its contributions
 * in list order.  The table's size and layout do not change any value.
 *
 * Rows with at most `cap` products and at most `lmax` A entries take this
 * order; the others take BRMerge-Precise's tree (merge_row), as the
 * library's hash tier does.  cap < 0: every row; lmax < 0: no list limit.  Output as brmo_brmerge_precise (int64 row pointer).
 */
typedef struct { int c; double v; } brmo_cv;
static int brmo_cv_less(const void *x, const void *y) {
    const int a = ((const brmo_cv *)x)->c, b = ((const brmo_cv *)y)->c;
    return (a > b) - (a < b);
}

int brmo_hash_numeric(const int *arpt, const int *acol, const double *aval,
                      const int *brpt, const int *bcol, const double *bval,
                      int M, int K, int N, int64_t cap, int lmax,
                      int64_t **crpt_, int **ccol_, double **cval_, int64_t *cnnz_, int nthreads) {
    (void)K;
    int *ccol = NULL;
    double *cval = NULL;
    int64_t *crpt = NULL;
    /* the structure (and the values of rows above cap) from Precise */
    int64_t *pcrpt = NULL;
    int *pccol = NULL;
    double *pcval = NULL;
    int64_t pnnz = 0;
    if (brmo_brmerge_precise(arpt, acol, aval, brpt, bcol, bval, M, K, N, &pcrpt, &pccol, &pcval, &pnnz, nthreads,
                             NULL) != 0)
        return -1;
    crpt = pcrpt;
    ccol = pccol;
    cval = pcval;
    int fail = 0;
    const int T = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(T)
    {
        int hcap = 0;
        int *htk = NULL;
        double *htv = NULL;
        brmo_cv *pv = NULL;
        int pvcap = 0;
#pragma omp for schedule(dynamic, 64)
        for (int row = 0; row < M; row++) {
            int64_t ip = 0;
            for (int j = arpt[row]; j < arpt[row + 1]; j++) ip += brpt[acol[j] + 1] - brpt[acol[j]];
            const int L = arpt[row + 1] - arpt[row];
            if (ip == 0 || (cap >= 0 && ip > cap) || (lmax >= 0 && L > lmax)) continue;
            int64_t nz = ip < N ? ip : N; /* BIN.h:164 */
            int bid = 0;
            while (nz > ((int64_t)8 << bid)) bid++;
            const int hs = 8 << bid; /* MIN_HT_N << (bin_id - 1), bin_id = bid + 1 */
            if (hs > hcap) {
                free(htk); free(htv);
                htk = (int *)malloc(sizeof(int) * (size_t)hs);
                htv = (double *)malloc(sizeof(double) * (size_t)hs);
                hcap = hs;
                if (!htk || !htv) {
#pragma omp atomic write
                    fail = 1;
                    hcap = 0;
                    continue;
                }
            }
            for (int s = 0; s < hs; s++) htk[s] = -1;
            for (int j = arpt[row]; j < arpt[row + 1]; j++) {
                const double av = aval[j];
                for (int k = brpt[acol[j]]; k < brpt[acol[j] + 1]; k++) {
                    const double t = av * bval[k];
                    const int key = bcol[k];
                    unsigned h = ((unsigned)key * BRMO_HASH_SCALE) & (unsigned)(hs - 1);
                    for (;;) {
                        if (htk[h] == key) { htv[h] = t + htv[h]; break; }
                        if (htk[h] == -1) { htk[h] = key; htv[h] = t; break; }
                        h = (h + 1) & (unsigned)(hs - 1);
                    }
                }
            }
            const int64_t n = crpt[row + 1] - crpt[row];
            if (n > pvcap) {
                free(pv);
                pv = (brmo_cv *)malloc(sizeof(brmo_cv) * (size_t)n);
                pvcap = (int)n;
                if (!pv) {
#pragma omp atomic write
                    fail = 1;
                    pvcap = 0;
                    continue;
                }
            }
            int64_t m = 0;
            for (int s = 0; s < hs; s++)
                if (htk[s] != -1 && m < n) { pv[m].c = htk[s]; pv[m].v = htv[s]; m++; }
            if (m != n) {
#pragma omp atomic write
                fail = 2;
                continue;
            }
            qsort(pv, (size_t)m, sizeof(brmo_cv), brmo_cv_less);
            for (int64_t q = 0; q < m; q++) {
                ccol[crpt[row] + q] = pv[q].c;
                cval[crpt[row] + q] = pv[q].v;
            }
        }
        free(htk); free(htv); free(pv);
    }
    if (fail) { free(crpt); free(ccol); free(cval); return -1; }
    *crpt_ = crpt; *ccol_ = ccol; *cval_ = cval; *cnnz_ = pnnz;
    return 0;
}
