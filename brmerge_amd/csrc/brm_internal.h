// Internal declarations shared by the BRMerge HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "../../include/brmerge.h"

namespace brm {

// ------------------------------------------------------------------ tiers
// Row bins by intermediate-product count (ip) and list count (L = nnz of the A row).
enum RowBin : int {
    BIN_EMPTY = 0,   // ip == 0: empty C row
    BIN_SHORT = 1,   // ip <= 64: one wavefront per row, merge in registers
    BIN_MEDIUM = 2,  // ip <= MED_CAP and L <= MED_LMAX: one workgroup per row, merge in LDS
    BIN_LONG = 3,    // otherwise: LDS leaf subtrees + HBM-streamed upper tree levels
    BIN_WMED = 4,    // ip <= min(wmed_cap, WMED_SPLIT0) and L <= WMED_LMAX: one wavefront per row
    BIN_TINY = 5,    // ip <= TINY_MAX: two rows per wavefront, merge in registers
    BIN_WMED2 = 6,   // WMED_SPLIT0 < ip <= min(wmed_cap, WMED_SPLIT), L <= WMED_LMAX
    BIN_WMED3 = 7,   // WMED_SPLIT < ip <= min(wmed_cap, WMED_SPLIT2), L <= WMED_LMAX
    BIN_HASH = 8,    // ip <= hash_cap (> 0), ip and L <= HASH_SMALL: Hash-SpGEMM comparator tier (hash.hip)
    BIN_HASH2 = 9,   // other rows of ip <= hash_cap and L <= HASH_CAP: the same tier with a larger table
    BIN_WMED4 = 10,  // WMED_SPLIT2 < ip <= wmed_cap, L <= WMED_LMAX (register tier with bitmap ranks only)
    BIN_COUNT = 11
};

constexpr int SHORT_MAX = 64;
// Largest product count of a row in the Hash-SpGEMM tier, and of its
// small-table launch.
constexpr int HASH_CAP = 1024;
constexpr int HASH_SMALL = 128;
constexpr int TINY_MAX = 32;

// Wavefront-sized LDS merges for small medium rows (several rows resident
// per CU instead of one 256-thread workgroup per row).
constexpr int WMED_LMAX = 64;
constexpr int WMED_CAP = 1024;
// The register tier with bitmap ranks (B of at most 2^20 columns) takes rows
// of up to WMED_CAP_BITMAP products.
constexpr int WMED_CAP_BITMAP = 2048;
// The wavefront tier runs its rows in up to four launches split at these
// product counts, so that the smaller rows get the smaller kernel shapes.
constexpr int WMED_SPLIT0 = 256;
constexpr int WMED_SPLIT = 512;
constexpr int WMED_SPLIT2 = 1024;

// LDS tree-merge configurations (threads, element capacity, list capacity).
constexpr int MED_NT = 256;
constexpr int MED_CAP = 4096;
constexpr int MED_LMAX = 512;
constexpr int MED_LOG_LMAX = 9;

// Long-row group merge: the tree above the LDS frontier is merged PASS_BITS
// levels per HBM pass; one workgroup streams up to GROUP_KMAX child lists
// through an LDS window of GROUP_CAP elements.
constexpr int PASS_BITS = 5;
constexpr int GROUP_KMAX = 1 << PASS_BITS;
constexpr int MAX_PASSES = 8;
// The bitmap-rank group merge (bmerge.hip) takes up to 2^BPASS_BITS children.
constexpr int BPASS_BITS = 6;
constexpr int BGROUP_KMAX = 1 << BPASS_BITS;
constexpr int GROUP_CFG_BITMAP = 4;

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

// Child list of a group merge: its source and its node id (first list
// position relative to the group's subtree root).
struct Child {
    Src src;
    int lid;
    int pad;
};

// One node above the frontier, merged from its children in one pass.
struct GroupTask {
    int row;
    int nchild;
    int final_dst;    // 1: writes C (Precise) / C-bar (Upper); 0: workspace
    int dst_buf;
    int64_t cbase;    // first child in the child array
    int64_t dst_off;  // element offset in destination
    int64_t slot;     // workspace: length slot of this node
};

struct LeafTask {
    int row;
    int ab, ae;       // A-entry range (an aligned subtree of the row's lists)
    int buf;
    int final_dst;    // the subtree is the whole row: write the destination
    int pad;
    int64_t dst_off;
    int64_t slot;
};

// Counters written by the planner and the merge kernels.
constexpr int CNT_LEAF = 0;      // leaf tasks
constexpr int CNT_MAXP = 1;      // passes needed by the wave
constexpr int CNT_LEAF_IP = 2;   // products inside leaves
constexpr int CNT_LEAF_OUT = 3;  // leaf output elements
constexpr int CNT_GRP_IN = 4;    // group merge input elements
constexpr int CNT_GRP_OUT = 5;   // group merge output elements
constexpr int CNT_CHILD = 6;     // child descriptors used
constexpr int CNT_ROUNDS = 7;    // streaming rounds of the group merges
constexpr int CNT_TASK0 = 8;     // CNT_TASK0 + j: group tasks of pass j (1-based)
constexpr int CNT_PH0 = CNT_TASK0 + MAX_PASSES + 1;  // bitmap engine: shader cycles per round phase
constexpr int N_PHASES = 8;
constexpr int CNT_N = CNT_PH0 + N_PHASES;

struct LongArgs {
    const int* rows;        // long rows of this wave
    int nrows;
    const int64_t* ws_base;  // per row: [2] element offsets into workspace buffers 0 and 1
    const int64_t* slot_base;
    const int* arpt;
    const int* acol;
    const double* aval;
    const int* brpt;
    const int64_t* ipA;
    const int64_t* final_off;  // per row: C offset (Precise: crpt) or nullptr (Upper: ipA[arpt[row]])
    int64_t* counters;
    LeafTask* leaves;
    int64_t leaves_cap;
    GroupTask* tasks;           // segmented by pass
    uint32_t* task_key;         // beside tasks: the node's product count (saturated), or nullptr
    const int64_t* task_seg;    // [MAX_PASSES+2] segment start per pass
    Child* children;
    int64_t child_cap;
    int frontier_cap; // frontier: products per LDS subtree
    int lmax;         // frontier: lists per LDS subtree (a power of two)
    int log_lmax;
    int pass_bits;    // tree levels merged per group pass
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
// Growable device arrays (virtual address reservation, physical memory mapped
// as they grow); released by dev_free(base).  vm_reserve returns nullptr when
// the device or driver does not support it.
void* vm_reserve(size_t bytes, int dev);
bool vm_grow(void* base, size_t bytes);  // maps [0, bytes); false on failure
void dev_trim();                 // release cached blocks to the driver
size_t dev_cached_bytes();
void dev_set_fail_after(int64_t n);  // test hook: fail the n-th next allocation (-1 = off)
// While an arena is set, the calling thread's allocations are carved from
// [base, base + bytes) (256-byte aligned, never released, nullptr when it is
// exhausted) instead of the caching allocator.
void arena_begin(void* base, size_t bytes);
void arena_end();
int64_t arena_remaining();  // -1: no arena
size_t arena_used();

// ------------------------------------------------------------------ tunables
struct Params {
    int64_t ws_budget_bytes;   // long-row workspace budget per wave (0 = auto)
    int med_cap;               // <= MED_CAP
    int force_long;            // test hook: route every non-empty row to the long tier
    int leaf_cap;              // frontier capacity for long-row leaves (<= MED_CAP)
    int profile;               // bracket every kernel launch with hipEvents
    int lds_cfg;               // LDS tier shape (threads, elements, lists): see lds_shape()
    int group_cfg;             // group-merge shape (threads, window elements): see group_shape()
    int wmed_cap;              // wavefront tier: products per row (0 = off, <= WMED_CAP_BITMAP, -1 = by N)
    int64_t launch_chunk;      // test hook: rows/tasks per launch (0 = the grid limit)
    int wmed_engine;           // wavefront tier numeric: 0 LDS merge levels, 1 register bitonic sort
    int64_t cbar_budget_bytes; // Upper: C-bar bytes per row wave (0 = auto)
    int64_t sym_chunk_products; // long-row symbolic: products per list chunk (1 .. 2^30, default 2^30)
    int hash_cap;              // rows of at most hash_cap products take the hash tier (0 = off, <= HASH_CAP)
    int upper_vm;              // Upper waves grow C in reserved address ranges (1) or copy chunks (0)
    int small_fused;           // products of few short rows in three launches (1) or the general pipeline (0)
    int wmed_count;            // wavefront tier symbolic: 1 column bitmap (B of <= 2^20 columns), 0 hash
    int wmed_bitmap;           // register tier numeric: 1 bitmap ranks (B of <= 2^20 columns, rows > 128 products), 0 sort
};

// Shapes of the LDS tree-merge kernels, selectable at run time.
struct LdsShape {
    int nt, cap, lmax, log_lmax;
};
constexpr int N_LDS_CFG = 3;
inline LdsShape lds_shape(int cfg) {
    switch (cfg) {
        case 1: return {256, 2048, 256, 8};
        case 2: return {128, 2048, 256, 8};
        default: return {256, 4096, 512, 9};
    }
}
constexpr int N_GROUP_CFG = 5;
Params& params();

// ------------------------------------------------------------------ launch limits
// HSA dispatch grids count work-items in 32 bits: launches with one
// workgroup (or wavefront) per row or task are split into chunks of at most
// 2^31 work-items (fewer under the launch_chunk test hook).
inline int64_t launch_chunk(int threads_per_item) {
    const int64_t lim = ((int64_t)1 << 31) / threads_per_item;
    const int64_t o = params().launch_chunk;
    return (o > 0 && o < lim) ? o : lim;
}

// ------------------------------------------------------------------ kernel attributes
// Raises one kernel's dynamic LDS limit the first time any thread launches
// it (worker threads of the request server may race to the first launch).
// Declare one `static SmemLimit` per kernel instantiation.
struct SmemLimit {
    std::once_flag once;
    hipError_t err = hipSuccess;
    hipError_t raise(const void* kern, int bytes) {
        std::call_once(once, [&] { err = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, bytes); });
        return err;
    }
};

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
// Product count of each A entry (|B row acol[j]|, 0 for a column outside [0, K)).
hipError_t scan_alen(const int* acol, const int* brpt, int K, int64_t n, int64_t* out, hipStream_t s);
hipError_t scan_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s);
hipError_t scan_i32(const int* in, int64_t n, int64_t* out, hipStream_t s);

}  // namespace brm
