// Internal declarations shared by the BRMerge HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/brmerge.h"

namespace brm {

// ------------------------------------------------------------------ tiers
// Row bins by intermediate-product count (ip) and list count (L = nnz of the A row).
enum RowBin : int {
    BIN_EMPTY = 0,   // ip == 0: empty C row
    BIN_SHORT = 1,   // ip <= 64: one wavefront per row, merge in registers
    BIN_MEDIUM = 2,  // ip <= MED_CAP and L <= MED_LMAX: one workgroup per row, merge in LDS
    BIN_LONG = 3,    // otherwise: LDS leaf subtrees + HBM-streamed upper tree levels
    BIN_COUNT = 4
};

constexpr int SHORT_MAX = 64;

// LDS tree-merge configurations (threads, element capacity, list capacity).
constexpr int MED_NT = 256;
constexpr int MED_CAP = 4096;
constexpr int MED_LMAX = 512;
constexpr int MED_LOG_LMAX = 9;

// Upper-level merge tiles (threads x items).
constexpr int LVL_NT = 256;
constexpr int LVL_ITEMS = 8;
constexpr int LVL_TILE = LVL_NT * LVL_ITEMS;

constexpr int MAX_HEIGHT = 32;

// ------------------------------------------------------------------ long-row plan
enum SrcKind : int { SRC_EMPTY = 0, SRC_VIRTUAL = 1, SRC_STORED = 2 };

// One input list of an upper-level merge.  VIRTUAL lists are a single scaled
// B row read in place (bcol/bval at `off`, values scaled by `scale`); STORED
// lists live in workspace buffer `buf` at element offset `off`, with their
// length written at runtime into slots[`slot`].
struct Src {
    int kind;
    int buf;
    int64_t off;
    int64_t len;   // VIRTUAL: length; STORED: unused (read from slot)
    int64_t slot;  // STORED: length slot
    double scale;  // VIRTUAL: A value
};

struct PairTask {
    Src L, R;
    int row;
    int final_dst;    // 1: writes C (Precise) / C-bar (Upper); 0: workspace
    int dst_buf;
    int pad;
    int64_t dst_off;  // element offset in destination
    int64_t slot;     // workspace: length slot of this node
};

struct LeafTask {
    int row;
    int ab, ae;       // A-entry range (an aligned subtree of the row's lists)
    int buf;
    int64_t dst_off;
    int64_t slot;
};

// Counters written by the planner.
constexpr int CNT_LEAF = 0;
constexpr int CNT_MAXH = 1;
constexpr int CNT_LEAF_IP = 2;
constexpr int CNT_LEAF_OUT = 3;
constexpr int CNT_TASK0 = 8;  // CNT_TASK0 + h: pair tasks at height h
constexpr int CNT_N = CNT_TASK0 + MAX_HEIGHT + 1;

struct LongArgs {
    const int* rows;        // long rows of this wave
    int nrows;
    const int64_t* ws_base;  // per row: element offset into workspace buffers
    const int64_t* slot_base;
    const int* arpt;
    const int* acol;
    const double* aval;
    const int* brpt;
    const int64_t* ipA;
    const int64_t* final_off;  // per row: C offset (Precise: crpt) or nullptr (Upper: ipA[arpt[row]])
    int64_t* counters;
    LeafTask* leaves;
    PairTask* tasks;            // segmented by height
    const int64_t* task_seg;    // [MAX_HEIGHT+1] segment start per height
};

// ------------------------------------------------------------------ error flags (device)
enum DevErr : int {
    DERR_NONE = 0,
    DERR_COUNT_MISMATCH = 1,   // numeric length != symbolic count
    DERR_CAPACITY = 2,         // planner/tile capacity exceeded
    DERR_INVALID = 4,          // input validation failed
};

// ------------------------------------------------------------------ allocator
void* dev_alloc(size_t bytes);   // nullptr on failure
void dev_free(void* p);
// Release after the calling thread's current library call has synchronised
// its stream (blocks still referenced by queued kernels).
void dev_defer_free(void* p);
void dev_flush_deferred();
void dev_trim();                 // release cached blocks to the driver
size_t dev_cached_bytes();
void dev_set_fail_after(int64_t n);  // test hook: fail the n-th next allocation (-1 = off)

// ------------------------------------------------------------------ tunables
struct Params {
    int64_t ws_budget_bytes;   // long-row workspace budget per wave (0 = auto)
    int med_cap;               // <= MED_CAP
    int force_long;            // test hook: route every non-empty row to the long tier
    int leaf_cap;              // frontier capacity for long-row leaves (<= MED_CAP)
    int profile;               // bracket every kernel launch with hipEvents
};
Params& params();

// ------------------------------------------------------------------ profiling
void prof_begin(const char* name, hipStream_t s);
void prof_end(hipStream_t s);
struct ProfScope {
    hipStream_t s;
    ProfScope(const char* name, hipStream_t st) : s(st) { prof_begin(name, st); }
    ~ProfScope() { prof_end(s); }
};

// ------------------------------------------------------------------ launchers
// Exclusive scans: out[0..n] with out[n] = total.
hipError_t scan_alen(const int* acol, const int* brpt, int64_t n, int64_t* out, hipStream_t s);
hipError_t scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s);
hipError_t scan_i32(const int* in, int64_t n, int64_t* out, hipStream_t s);

}  // namespace brm
