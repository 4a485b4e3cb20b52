// Kernel argument blocks and launcher declarations shared between the
// kernel translation units and the host orchestration (api.hip).
#pragma once
#include "brm_internal.h"

namespace brm {

// Read-only view of A and B (device pointers).
struct RowArgs {
    const int* arpt;
    const int* acol;
    const double* aval;
    const int* brpt;
    const int* bcol;
    const double* bval;
    const int64_t* ipA;  // [nnzA+1]: global index of each A entry's first product
    int M, K, N;
};

// Where finished rows go.
struct OutArgs {
    int upper;            // 0: Precise (C at crpt), 1: Upper (C-bar at ipA[arpt[row]], lengths to rowlen)
    const int64_t* crpt;  // Precise
    int* ccol;            // Precise: C; Upper: C-bar
    double* cval;
    int64_t* rowlen;      // Upper
    int* err;             // DevErr bits
    int slot;             // Upper, short tiers: row k of the row list writes at k * slot (0: at ipA)
};

// flop.hip
hipError_t launch_validate(const RowArgs& a, int64_t nnzA, int64_t nnzB, int* err, hipStream_t s);
// nnz = rpt[n] (checked by launch_validate before; not followed when *err is set)
hipError_t launch_validate_sorted(const int* rpt, const int* col, int n, int64_t nnz, int* err, hipStream_t s);
// bins: [BIN_COUNT-1][M] row lists (bin b at (b-1)*M); bin_cnt: [BIN_COUNT]
// Rows are all treated as empty when *err is set (err may be null).
hipError_t launch_rowstats(const RowArgs& a, int med_cap, int med_lmax, int wmed_cap, int force_long, int hash_cap,
                           int64_t* iprow, int64_t* nnzrow, int* bins, int* bin_cnt, const int* err, hipStream_t s);

// short.hip
// group: lanes per row, 64 (ip <= 64) or 32 (ip <= 32, two rows per wavefront)
hipError_t launch_short_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s,
                              int group = 64);
hipError_t launch_short_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                                int group = 64);
// C rows of short rows from their slots (row k of the list at k * group)
hipError_t launch_short_copy(const int* rows, int n, const int* scol, const double* sval, const int64_t* crpt,
                             int* ccol, double* cval, hipStream_t s, int group);

// lds.hip
hipError_t launch_medium_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s);
hipError_t launch_medium_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                                 int lds_cfg);
hipError_t launch_long_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s);
hipError_t launch_wmed_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s, int wmed_cap);
hipError_t launch_wmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s,
                               int wmed_cap);
hipError_t launch_leaf(const LeafTask* tasks, int64_t n, const RowArgs& a, int* const ws_col[2],
                       double* const ws_val[2], int64_t* slots, int64_t* counters, const OutArgs& o,
                       hipStream_t s);

// long.hip
hipError_t launch_plan(const LongArgs& la, int* err, hipStream_t s);
// group_cfg: the engine the tasks were planned for (GROUP_CFG_BITMAP: bmerge.hip)
// order: task indices to run, in that order (workgroup b runs tasks[order[b]]), or nullptr
// for tasks[b]; the bitmap engine only
hipError_t launch_group(const GroupTask* tasks, int64_t n, const Child* children, const RowArgs& a,
                        int* const ws_col[2], double* const ws_val[2], int64_t* slots, int64_t* counters,
                        const OutArgs& o, hipStream_t s, int group_cfg, const int* order);
// Task order of one pass, largest node first: idx_out = the indices 0..n-1
// sorted by keys descending (keys_out receives the sorted keys).  tmp of
// order_tasks_tmp_bytes(n) bytes.
size_t order_tasks_tmp_bytes(int n);
// Rows sorted in place by product count, largest first (keys, keys_out and
// rows_out: n entries each; tmp of order_rows_tmp_bytes(n) bytes).
size_t order_rows_tmp_bytes(int n);
hipError_t order_rows(int* rows, int n, const int* arpt, const int64_t* ipA, uint64_t* keys, uint64_t* keys_out,
                      int* rows_out, void* tmp, size_t tmp_bytes, hipStream_t s);
hipError_t order_tasks(const uint32_t* keys, int n, uint32_t* keys_out, int* idx_in, int* idx_out, void* tmp,
                       size_t tmp_bytes, hipStream_t s);

// rmed.hip (register tier; cap = products per row, <= WMED_CAP)
hipError_t launch_rmed_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap);
// Exact row lengths of register-tier rows (ip <= cap, at most 64 lists) by
// an LDS column bitmap; only for B of at most 2^20 columns (rmed_count_fits).
bool rmed_count_fits(int N);
// Largest register-tier row (products) for B of N columns under the current
// knobs: WMED_CAP_BITMAP where the bitmap ranks apply, else WMED_CAP.
int rmed_max_cap(int N);
hipError_t launch_rmed_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s, int cap);
// Wavefront-tier capacity chosen when wmed_cap is -1 (register engine): 1024
// products while 32-bit sort keys hold N columns, else 512.
int rmed_auto_cap(int N);

// small.hip: products of at most SMALL_MAX_ROWS rows, all in the short tier.
constexpr int SMALL_MAX_ROWS = 1 << 20;
constexpr int SMALL_SLOTS = 32;  // partial-sum counters (summed by the host)
struct SmallState {
    int err;     // DevErr bits
    int miss;    // a row has more than G lists or products: nothing of C is usable
    int ticket;  // workgroup order of k_small
    int pad;
    int64_t nnz[SMALL_SLOTS];  // nnz(C), in partial sums
    int64_t ip[SMALL_SLOTS];   // intermediate products, in partial sums
};
// int64 words of the state block (SmallState, then the look-back states)
// for M rows in groups of `group` (32 or 64) lanes
int64_t small_state_words(int M, int group);
// Clears the state block, then merges every row and writes C's row pointer
// and C (capacity group * M entries).
hipError_t launch_small(const RowArgs& a, int64_t nnzA, int64_t nnzB, int64_t* crpt, int* ccol, double* cval,
                        SmallState* st, int group, hipStream_t s);

// hash.hip (rows of at most HASH_CAP products)
// cap: the largest product count of the listed rows (<= HASH_SMALL picks the small table)
hipError_t launch_hash_numeric(const int* rows, int n, const RowArgs& a, const OutArgs& o, hipStream_t s, int cap);
// lds.hip: exact row counts of hash-tier rows
hipError_t launch_hash_count(const int* rows, int n, const RowArgs& a, int64_t* nnzrow, hipStream_t s, int cap);

// bmerge.hip
hipError_t launch_bgroup(const GroupTask* tasks, int64_t n, const Child* children, const RowArgs& a,
                         int* const ws_col[2], double* const ws_val[2], int64_t* slots, int64_t* counters,
                         const OutArgs& o, hipStream_t s, const int* order);

// assemble.hip
hipError_t launch_compact(const int64_t* crpt, int M, const RowArgs& a, const int* cbar_col,
                          const double* cbar_val, int* ccol, double* cval, int64_t nnzC, hipStream_t s);
hipError_t launch_rpt_to_i32(const int64_t* in, int64_t n, int* out, hipStream_t s);
// *out = order-independent 64-bit checksum of nwords 4-byte words at p
hipError_t launch_checksum(const void* p, int64_t nwords, uint64_t* out, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);
hipError_t launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s);
hipError_t launch_fill_i32(int* p, int64_t n, int v, hipStream_t s);
hipError_t launch_long_info(const int* rows, int n, const RowArgs& a, int64_t* ip_out, int* len_out,
                            hipStream_t s);

}  // namespace brm
