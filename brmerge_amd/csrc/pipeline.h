// Host-side pipeline pieces shared by the single-product entries (api.hip),
// the batched and planned entries (batch.hip) and the request server
// (serve.hip).
#pragma once
#include <initializer_list>
#include <algorithm>
#include <string>
#include <vector>

#include "kernels.h"

namespace brm {

void set_err(const char* fmt, ...);
int hip_status(hipError_t e, const char* what);
int use_device();
int d2h(void* dst, const void* src, size_t bytes, hipStream_t s);
// several device-to-host copies behind one synchronisation
struct D2HPart {
    void* dst;
    const void* src;
    size_t bytes;
};
int d2h_parts(std::initializer_list<D2HPart> parts, hipStream_t s);
int check_dev_err(const int* derr, hipStream_t s, const char* stage);
int dev_err_status(int err_flags, const char* stage);
double now_s();
brm_stats& thread_stats();

#define HIPC(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_status(e_, #expr); \
    } while (0)

#define BRMC(expr)                     \
    do {                               \
        int st_ = (expr);              \
        if (st_ != BRM_OK) return st_; \
    } while (0)

// Device allocations owned by one call; freed on scope exit unless released.
struct Scratch {
    std::vector<void*> bufs;
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch() {
        for (void* p : bufs) dev_free(p);
    }
    template <typename T>
    int alloc(T** p, int64_t n, const char* what) {
        *p = (T*)dev_alloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
        if (!*p) {
            set_err("device allocation of %s (%lld elements) failed", what, (long long)n);
            return BRM_E_NOMEM;
        }
        bufs.push_back(*p);
        return BRM_OK;
    }
    void release(void* p) {
        auto it = std::find(bufs.begin(), bufs.end(), p);
        if (it != bufs.end()) bufs.erase(it);
    }
    void free_now(void* p) {
        release(p);
        dev_free(p);
    }
};

enum Phase { P_PRE, P_FLOP, P_PFLOP, P_LB, P_SYM, P_PNNZ, P_ALLOC, P_COMPUTE, P_COPY, P_CLEAN, P_N };

// hipEvent pairs around phases, summed per phase.
// A clock constructed with on = false records nothing (no events created).
struct PhaseClock {
    hipStream_t s;
    bool on;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    int open_phase = -1;
    hipEvent_t open_ev = nullptr;
    explicit PhaseClock(hipStream_t st, bool on_ = true) : s(st), on(on_) {}
    PhaseClock(const PhaseClock&) = delete;
    PhaseClock& operator=(const PhaseClock&) = delete;
    ~PhaseClock();
    void begin(int p);
    void end();
    void collect(double* out);
};

// A and B of one product, device pointers.
struct DevIn {
    const int* arpt;
    const int* acol;
    const double* aval;
    int64_t nnzA;
    const int* brpt;
    const int* bcol;
    const double* bval;
    int64_t nnzB;
    int M, K, N;
    int b_checked = 0;  // B already validated (row pointers, column range, sorted rows)
};

struct DevOut {
    int64_t* crpt;
    int* ccol;
    double* cval;
    int64_t nnz;
};

// The structure of one product: per-entry product offsets, row bins, and
// (after pattern_symbolic) the exact row lengths and row pointer of C.
// Values of A and B are read only by the numeric phase, so one Pattern
// serves any number of numeric passes over the same sparsity patterns.
struct Pattern {
    Scratch sc;
    RowArgs a{};
    int64_t nnzA = 0, nnzB = 0;
    int* derr = nullptr;
    int* bin_cnt = nullptr;
    int64_t* ipA = nullptr;
    int64_t* iprow = nullptr;
    int64_t* rowcnt = nullptr;
    int64_t* crpt = nullptr;
    int* bins = nullptr;  // [BIN_COUNT-1][M]: row lists of the non-empty bins
    int64_t* crpt_ext = nullptr;  // caller's row pointer of C (set before pattern_prepare), else allocated
    int bc[BIN_COUNT] = {};
    int64_t total_ip = 0;
    int64_t nnzC = 0;
    int lds_cfg = 0;
    int med_cap = MED_CAP;
    int wmed_cap = 0;
    // Short rows merged once, in the symbolic phase (pattern_symbolic with
    // stash_short): row k of the tiny / short list at k * 32 / k * 64 here
    bool stashed = false;
    int* stash_col[2] = {};
    double* stash_val[2] = {};
    const int* rows(int bin) const { return bins + (int64_t)(bin - 1) * a.M; }
};

// Largest stash of merged short rows (12 bytes per slot).
constexpr int64_t SHORT_STASH_MAX_BYTES = (int64_t)8 << 30;

// Validation, product offsets (flop scan) and row bins under the current
// tier parameters.
int pattern_prepare(const DevIn& in, hipStream_t s, PhaseClock* clk, Pattern& p);
// Re-bins rows for LDS shape `lds_cfg`, medium capacity `med_cap` and
// wavefront-tier capacity `wmed_cap`.
int pattern_bin(Pattern& p, int lds_cfg, int med_cap, int wmed_cap, int force_long, hipStream_t s);
// Exact nnz per row, C's row pointer and nnz(C) (BRMerge-Precise symbolic).
// With stash_short, rows of at most 64 products are merged here already
// (their lengths are the counts) when the stash fits SHORT_STASH_MAX_BYTES;
// the numeric phase then copies them.  A Pattern whose values change
// between numeric passes must not stash.
// With read_nnz false, nnz(C) stays on the device (p.nnzC = -1).
int pattern_symbolic(Pattern& p, hipStream_t s, PhaseClock* clk, bool stash_short = false, bool read_nnz = true);
// Numeric phase of BRMerge-Precise into C (ccol/cval of nnz(C) elements).
// The short and medium tiers only enqueue kernels (capturable in a graph);
// the long tier needs host round trips between its launches.
int pattern_numeric_tiers(const Pattern& p, int* ccol, double* cval, hipStream_t s);
int pattern_numeric_long(const Pattern& p, int* ccol, double* cval, hipStream_t s);

// Caller-owned C of a Precise product: row pointer (M+1) and cap entries of
// columns and values.  When nnz(C) > cap the call fails with BRM_E_OVERFLOW
// and reports nnz(C) in *nnz_needed.
struct ExtOut {
    int64_t* crpt;
    int* ccol;
    double* cval;
    int64_t cap;
    int64_t* nnz_needed;
};

// One SpGEMM.  ph (phase seconds, P_N entries) may be null: then no phase
// events are recorded.  ext (Precise only) may be null: then C is allocated.
int run_spgemm(const DevIn& in, bool upper, int64_t max_nnz, hipStream_t s, DevOut* out, double* ph,
               const ExtOut* ext = nullptr);

// The calling thread's device for library calls (set by brm_init).
int thread_device();
void set_thread_device(int device);
// Long-row workspace budget of the calling thread's calls (0: read the
// device's free memory at each call).
int64_t thread_ws_budget();
void set_thread_ws_budget(int64_t bytes);
// Whether the calling thread's calls may put work on side streams beside
// their own (the request server's workers do not: several workers with two
// side streams each would outnumber the GPU's hardware queues, and streams
// sharing a queue wait on each other).
bool thread_side_streams();
void set_thread_side_streams(bool on);
// Device address of the no-allocation entry's fixed scratch on `device`
// (the current device), resolved once per device.
int noalloc_setup(int device);

// Fused operands of a batch: A_i stacked by rows with columns shifted to
// their B_i's rows; B_i stacked by rows, or, when a_col_off is given, not
// assembled at all (the A columns are then shifted by a_col_off[i]).
struct FusedOut {
    Scratch sc;
    DevIn in{};
    std::vector<int64_t> mbase;  // first fused row of each A_i, [n+1]
};
int fuse_pairs(const brm_pair* pairs, int n, const int64_t* a_col_off, hipStream_t s, FusedOut* out);
// Row pointers restarting at 0 for each product of a fused result, and each
// product's first entry in the fused column/value arrays.
hipError_t launch_split_rpt(const int64_t* crpt, const int64_t* mbase_d, const int64_t* ebase_d, int n,
                            int64_t total, int64_t* rpt_out, int64_t* base_out, hipStream_t s);

}  // namespace brm
