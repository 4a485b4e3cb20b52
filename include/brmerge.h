/*
 * brmerge.h -- C ABI of the MI355X (gfx950) BRMerge SpGEMM library
 * (libbrmerge.so).  C = A * B for CSR matrices with int32 column indices and
 * FP64 values, computed by binary-row merging: every A entry of a row selects
 * one B row scaled by the A value, and the row's lists are merged pairwise in
 * an aligned binary tree.  Results are bit-identical to
 * BRMerge-SpGEMM/inc/spgemm.h (same tree bracketing, products rounded before
 * any add).
 *
 * Conventions shared by every entry point:
 *   - 0-based CSR; columns of each B row strictly increasing (the reference
 *     guarantees this at read time, src/CSR.cpp:454-477; here it is checked
 *     and rejected with BRM_E_INVALID);
 *   - K and N are the inner and output dimensions (the reference accepts and
 *     ignores them, spgemm.h:468-471; here they bound the index checks);
 *   - every call is synchronous and returns a BRM_* status; the message of
 *     the last failure is available from brm_last_error().
 */
#ifndef BRMERGE_H_
#define BRMERGE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BRM_OK 0
#define BRM_E_INVALID 1   /* malformed CSR or argument */
#define BRM_E_NOMEM 2     /* device or host allocation failed; no output is returned */
#define BRM_E_OVERFLOW 3  /* nnz(C) does not fit the int32 row pointer of an _i32 entry */
#define BRM_E_DEVICE 4    /* HIP runtime error */
#define BRM_E_INTERNAL 5  /* internal consistency check failed */

/* Phase times in seconds.  Field names follow Precise_Timing
 * (BRMerge-SpGEMM/inc/Timings.h:27-46); h2d/d2h are the host<->device copies
 * of the host-pointer entries (zero for device entries). */
typedef struct brm_precise_timing {
    double pre_allocate;
    double compute_flop;
    double prefix_sum_flop;
    double load_balance;
    double symbolic;
    double prefix_sum_nnz;
    double allocate_C;
    double compute;
    double cleanup;
    double total;
    double h2d;
    double d2h;
} brm_precise_timing;

/* Field names follow Upper_Timing (BRMerge-SpGEMM/inc/Timings.h:4-25). */
typedef struct brm_upper_timing {
    double thread;
    double pre_allocate;
    double compute_flop;
    double prefix_sum_flop;
    double load_balance;
    double allocate_hook;
    double compute;
    double prefix_sum_nnz;
    double allocate_C;
    double copy;
    double cleanup;
    double total;
    double h2d;
    double d2h;
} brm_upper_timing;

/* Select the HIP device used by this thread's calls (default 0). */
int brm_init(int device);
const char *brm_last_error(void);
const char *brm_version(void);

/*
 * Host-pointer entries.  Replace
 *   brmerge_precise(mint*,mint*,mdouble*,mint*,mint*,mdouble*,mint**,mint**,mdouble**,
 *                   mint M,mint K,mint N,mint* cnnz,Precise_Timing&)   spgemm.h:468-471
 *   brmerge_upper(... , Upper_Timing&)                                  spgemm.h:220-223
 * with the same argument order.  *crpt, *ccol and *cval are allocated with
 * C++ new[] (as the reference does, spgemm.h:479,574-575) and owned by the
 * caller: release them with delete[] (CSR::release, CSR.cpp:102-109) or with
 * brm_host_free_i32 / brm_host_free_f64.  On error nothing is allocated.
 * The timing pointer may be NULL.
 */
int brm_spgemm_precise_i32(const int *arpt, const int *acol, const double *aval,
                           const int *brpt, const int *bcol, const double *bval,
                           int **crpt, int **ccol, double **cval,
                           int M, int K, int N, int *cnnz, brm_precise_timing *timing);
int brm_spgemm_upper_i32(const int *arpt, const int *acol, const double *aval,
                         const int *brpt, const int *bcol, const double *bval,
                         int **crpt, int **ccol, double **cval,
                         int M, int K, int N, int *cnnz, brm_upper_timing *timing);

/* As above with an int64 row pointer for C (nnz(C) may exceed INT32_MAX;
 * the reference's int32 crpt cannot represent such products, spgemm.h:479,571).
 * Free *crpt with brm_host_free_i64. */
int brm_spgemm_precise_i64(const int *arpt, const int *acol, const double *aval,
                           const int *brpt, const int *bcol, const double *bval,
                           int64_t **crpt, int **ccol, double **cval,
                           int M, int K, int N, int64_t *cnnz, brm_precise_timing *timing);
int brm_spgemm_upper_i64(const int *arpt, const int *acol, const double *aval,
                         const int *brpt, const int *bcol, const double *bval,
                         int64_t **crpt, int **ccol, double **cval,
                         int M, int K, int N, int64_t *cnnz, brm_upper_timing *timing);

void brm_host_free_i32(int *p);
void brm_host_free_i64(int64_t *p);
void brm_host_free_f64(double *p);

/*
 * Device-pointer entries: A and B already in HBM (device pointers), C
 * returned in HBM.  `stream` is a hipStream_t (NULL = default stream); the
 * call returns after the stream has completed the product.  Output buffers
 * come from the library's caching device allocator: release each with
 * brm_dev_free.  nnzA / nnzB are arpt[M] / brpt[K] (passed to avoid a
 * device read).
 */
int brm_spgemm_precise_dev(const int *arpt, const int *acol, const double *aval, int64_t nnzA,
                           const int *brpt, const int *bcol, const double *bval, int64_t nnzB,
                           int M, int K, int N,
                           int64_t **crpt, int **ccol, double **cval, int64_t *cnnz,
                           brm_precise_timing *timing, void *stream);
int brm_spgemm_upper_dev(const int *arpt, const int *acol, const double *aval, int64_t nnzA,
                         const int *brpt, const int *bcol, const double *bval, int64_t nnzB,
                         int M, int K, int N,
                         int64_t **crpt, int **ccol, double **cval, int64_t *cnnz,
                         brm_upper_timing *timing, void *stream);

void *brm_dev_malloc(uint64_t bytes);  /* caching device allocator */
/* Device copy of an int64 row pointer (n entries) into int32, so that a
 * product from a _dev entry can be the operand of the next product (e.g.
 * P^T * (A * P)); returns BRM_E_OVERFLOW if a value exceeds INT32_MAX. */
int brm_rpt_to_i32(const int64_t *in, int64_t n, int *out, void *stream);
/* Transpose of an M x N device CSR (int32 indices) into caller-provided
 * device arrays trpt[N+1], tcol[nnz], tval[nnz]; rows of the transpose come
 * out sorted.  The reference transposes on the host by sorting
 * (Existing-SpGEMM/inc/convert.h:6-58); brm_csr_transpose is the host form. */
int brm_csr_transpose_dev(const int *rpt, const int *col, const double *val, int64_t nnz, int M, int N,
                          int *trpt, int *tcol, double *tval, void *stream);
void brm_dev_free(void *p);
void brm_dev_trim(void);               /* return cached device memory to the driver */

/*
 * Intermediate-product ("flop") count, the reference's
 *   long compute_flop(const CSR& A, const CSR& B)     spgemm.h:59-66
 * (sum over A entries of the selected B row's length) for host pointers.
 * Returns BRM_OK and the total in *total.  The _dev form also fills the
 * per-row counts (device int64[M]) when row_ip is non-NULL.
 */
int brm_flop_count(const int *arpt, const int *acol, const int *brpt, int M, int K, int64_t *total);
int brm_flop_count_dev(const int *arpt, const int *acol, int64_t nnzA, const int *brpt, int M, int K,
                       int64_t *row_ip, int64_t *total, void *stream);

/* Tuning knobs (names: "ws_budget_bytes", "med_cap", "leaf_cap",
 * "force_long", "profile", "lds_cfg", "group_cfg", "wmed_cap",
 * "wmed_engine", "launch_chunk", "cbar_budget_bytes", "sym_chunk_products",
 * "hash_cap", "upper_vm", "small_fused", "wmed_count", "wmed_bitmap");
 * returns BRM_E_INVALID for unknown names or
 * values.  Knobs other than hash_cap never change results, only speed.  group_cfg 0-3:
 * merge-path engine shapes for long rows; 4 (default): the bitmap-rank
 * engine.  wmed_cap: products per row of the wavefront tier (0 off, up to
 * 2048, above 1024 only where the register tier ranks with a bitmap; -1, the
 * default: 2048 when B has at most 2^20 columns, else 1024 when it has at
 * most 65,536, else 512).
 * cbar_budget_bytes: BRMerge-Upper's C-bar bytes per row wave (0, the
 * default: a quarter of the memory the call can get, at most an eighth of
 * the device).  sym_chunk_products: products per list chunk of the
 * long-row symbolic count (1 to 2^30, the default).  hash_cap (0, the
 * default, to 1024): rows of at most hash_cap products take the Hash-SpGEMM
 * tier instead of BRMerge's (the comparison baseline of the reference's
 * Existing-SpGEMM/inc/hash_mult.h:558-608: each column's contributions added
 * in list order); this knob, alone among them, changes value bits (within
 * 1e-12 of BRMerge's), never the structure.  upper_vm (1, the default):
 * BRMerge-Upper's row waves compact into C grown in a reserved address range;
 * 0 (or a driver without address reservation): into chunks copied into C.
 * small_fused (1, the default): products whose rows are all short run in
 * the three-launch fused path.  wmed_count (1, the default): the
 * wavefront tier's row lengths are counted with a column bitmap when B has
 * at most 2^20 columns; 0: with the hash count.  wmed_bitmap (1, the
 * default): the register tier ranks rows of more than 128 products with a
 * column bitmap when B has at most 2^20 columns (and then takes rows of up
 * to 2,048 products); 0: with a sort. */
int brm_set_param(const char *name, int64_t value);
int64_t brm_get_param(const char *name);

/* Synchronous copies between host memory and memory from brm_dev_malloc
 * (for callers without a HIP binding of their own). */
int brm_memcpy_h2d(void *dst_dev, const void *src_host, uint64_t bytes);
int brm_memcpy_d2h(void *dst_host, const void *src_dev, uint64_t bytes);

/* Per-kernel device timing.  With brm_set_param("profile", 1) every kernel
 * launch of the library is bracketed by a pair of hipEvents on its stream;
 * brm_prof_report writes {"kernel_name": [launches, total_ms], ...} as JSON
 * into buf (returns BRM_E_INVALID if cap is too small); brm_prof_reset
 * clears the totals. */
void brm_prof_reset(void);
int brm_prof_report(char *buf, uint64_t cap);

/* Work counters of the calling thread's last SpGEMM.  Used to turn kernel
 * times into algorithmic bandwidth. */
typedef struct brm_stats {
    int64_t rows_by_tier[4];   /* empty, short (<= 64 products), medium (LDS), long */
    int64_t total_ip;          /* intermediate products */
    int64_t nnz_c;
    int64_t long_waves;        /* workspace waves of the long-row tier */
    int64_t leaf_tasks;        /* LDS-merged aligned subtrees of long rows */
    int64_t leaf_ip;           /* products inside those subtrees */
    int64_t leaf_out;          /* merged elements they produce */
    int64_t passes;            /* group-merge passes (5 tree levels each) */
    int64_t group_tasks;       /* tree nodes merged above the subtrees */
    int64_t group_rounds;      /* LDS streaming rounds of those merges */
    int64_t group_in;          /* their input elements */
    int64_t group_out;         /* their output elements */
    int64_t upper_waves;       /* BRMerge-Upper: row waves of C-bar (1 when C-bar fits at once) */
} brm_stats;
int brm_last_stats(brm_stats *out);

/*
 * BRMerge-Precise with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller passes A and B (device) and C's arrays: crpt (M+1 entries) and
 * `cap` entries of ccol / cval.  Every scratch buffer comes from one fixed
 * device array of brm_noalloc_scratch_bytes() bytes (256 MiB), held by each
 * GPU the library runs on until the process ends.  Products whose scratch
 * would exceed it fail with BRM_E_NOMEM; when nnz(C) > cap the call fails
 * with BRM_E_OVERFLOW and *cnnz is nnz(C).  Calls on one GPU share the
 * array: the library serialises them, and each returns only once its
 * stream is idle.  Results are bit-identical to brm_spgemm_precise_dev.
 */
int brm_spgemm_precise_dev_noalloc(const int *arpt, const int *acol, const double *aval, int64_t nnzA,
                                   const int *brpt, const int *bcol, const double *bval, int64_t nnzB, int M, int K,
                                   int N, int64_t *crpt, int *ccol, double *cval, int64_t cap, int64_t *cnnz,
                                   void *stream);
uint64_t brm_noalloc_scratch_bytes(void);

/* Order-independent 64-bit checksum of `bytes` (a multiple of 4) of device
 * memory: the sum mod 2^64 of a SplitMix64 mix of (word index, word).  No
 * counterpart in the reference; it lets callers compare results of
 * billions of entries (Precise against Upper, one run against another)
 * without copying them to the host. */
int brm_checksum_dev(const void *p, uint64_t bytes, uint64_t *out, void *stream);

/*
 * Batched products.  The reference multiplies one pair per call
 * (brmerge_precise, spgemm.h:468-471); these entries run many independent
 * products C_i = A_i * B_i in one pass, as the block-diagonal product
 * diag(A_i) * diag(B_i).  Every C_i is bit-identical to a single call on
 * (A_i, B_i).  All pointers of a brm_pair are device pointers.
 */
typedef struct brm_pair {
    const int *arpt;
    const int *acol;
    const double *aval;
    int64_t nnzA;
    const int *brpt;
    const int *bcol;
    const double *bval;
    int64_t nnzB;
    int M, K, N;
} brm_pair;

/* One-shot batch (BRMerge-Precise).  C is returned as one device CSR holding
 * the rows of C_0, C_1, ... in order, with an int64 row pointer and each
 * product's own column indices; C_i is rows row_off[i] .. row_off[i+1]-1
 * (row_off: caller-provided host int64[npairs+1]).  Release the three
 * arrays with brm_dev_free.  Sums of rows, entries of A or entries of B over
 * the batch must fit int32. */
int brm_spgemm_batched_dev(const brm_pair *pairs, int npairs,
                           int64_t **crpt, int **ccol, double **cval, int64_t *cnnz,
                           int64_t *row_off, brm_precise_timing *timing, void *stream);

/* Batch plan for repeated products over fixed sparsity patterns (values of
 * A_i and B_i may change between executions; pointers may not).  Creation
 * builds the fused operands, runs the symbolic phase, allocates C and, when
 * `autotune` is non-zero, times the tier-shape candidates on this batch and
 * keeps the fastest.  The numeric phase (value refresh + short and medium
 * tiers) is captured in one HIP graph; execute replays it on `stream` (and
 * runs long rows, if any, eagerly after it) and waits for completion.
 * Result arrays stay owned by the plan until brm_batch_plan_destroy;
 * *row_off is a host array of npairs+1 entries. */
typedef struct brm_batch_plan brm_batch_plan;
int brm_batch_plan_create(const brm_pair *pairs, int npairs, int autotune,
                          brm_batch_plan **plan, void *stream);
int brm_batch_plan_execute(brm_batch_plan *plan, void *stream);
int brm_batch_plan_result(const brm_batch_plan *plan, const int64_t **crpt, const int **ccol,
                          const double **cval, int64_t *cnnz, const int64_t **row_off);
/* JSON summary: sizes, tier counts, chosen shape, tuner timings, the
 * batch's shape key and where the shape came from ("timed", "cache" or
 * "none" without autotune). */
int brm_batch_plan_info(const brm_batch_plan *plan, char *buf, uint64_t cap);
/* Tuned-shape cache (no counterpart in the reference): a JSON file of the
 * tier shapes the plan tuner chose, keyed by the batch's shape statistics
 * (log2 buckets of rows, intermediate products, columns and the row count of
 * every tier bin).  With a cache set, brm_batch_plan_create with `autotune`
 * first looks the batch's key up and, on a hit, uses that shape without
 * timing the candidates; on a miss it tunes and records the choice.  The
 * file is rewritten whole (temporary file, then rename).  NULL or "" turns
 * the cache off; until this is called, the BRM_TUNE_CACHE environment
 * variable names the file. */
int brm_set_tune_cache(const char *path);
void brm_batch_plan_destroy(brm_batch_plan *plan);

/*
 * Request server (no counterpart in the reference, which has no serving
 * path): B operands resident in HBM, C = A * B requests admitted against an
 * HBM budget by their BRMerge-Upper bound (12 bytes per intermediate product
 * plus 8 per row pointer entry) and fused into batches as they arrive;
 * `streams` worker threads run batches concurrently, one HIP stream each.
 * Each result is bit-identical to a single call on (A, B).
 *
 * brm_server_add_b copies B into the server (the caller's arrays may be
 * freed afterwards); it waits until no request is queued or running.  A of
 * a submitted request is read when its batch runs: keep it alive until the
 * request is done.  Results live in HBM until brm_server_release; their
 * bytes count against the budget until then.  Each worker's long-row
 * workspace is bounded by an eighth of the budget divided among the workers.
 * Times are seconds on the host's steady clock.
 */
typedef struct brm_server brm_server;
typedef struct brm_server_result {
    const int64_t *crpt;   /* device, M+1 entries, crpt[0] = 0 */
    const int *ccol;       /* device, nnz entries */
    const double *cval;    /* device, nnz entries */
    int64_t nnz;
    int64_t ip;            /* intermediate products of the request */
    int M, N;
    int status;            /* BRM_OK or the failure of the request's batch */
    int batch_size;        /* requests fused into the request's batch */
    double t_submit, t_start, t_done;
} brm_server_result;

int brm_server_create(int64_t budget_bytes, int max_batch, int streams, brm_server **srv);
int brm_server_add_b(brm_server *srv, const int *brpt, const int *bcol, const double *bval, int64_t nnzB,
                     int K, int N, int *b_id);
int brm_server_submit(brm_server *srv, int b_id, const int *arpt, const int *acol, const double *aval,
                      int64_t nnzA, int M, int64_t *req_id);
int brm_server_poll(brm_server *srv, int64_t req_id, int *done);
int brm_server_wait(brm_server *srv, int64_t req_id, brm_server_result *out);
int brm_server_release(brm_server *srv, int64_t req_id);
/* JSON: completed, batches, mean_batch, p50_s, p99_s, max_s, span_s, total_ip, in_use_bytes, queued. */
int brm_server_stats(brm_server *srv, char *buf, uint64_t cap);
/* Finishes queued requests, stops the workers and frees everything. */
void brm_server_destroy(brm_server *srv);

/* A plain streaming device copy of `bytes` (a multiple of 16) bytes, the
 * reference kernel against which measured HBM byte counters are scaled. */
int brm_bench_copy(void *dst, const void *src, uint64_t bytes, void *stream);

/* Test hook: make the n-th next device allocation fail (n < 0 disables). */
void brm_debug_fail_alloc_after(int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* BRMERGE_H_ */
