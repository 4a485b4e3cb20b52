// Host orchestration and the C ABI (include/brmerge.h).
//
// One SpGEMM call runs these phases on one HIP stream (field names follow
// Precise_Timing / Upper_Timing, BRMerge-SpGEMM/inc/Timings.h:4-46):
//   pre_allocate     scratch buffers, input validation
//   compute_flop     per-A-entry product counts + their prefix (one scan)
//   load_balance     per-row product counts and tier bins
//   symbolic         exact per-row nnz (Precise only)
//   prefix_sum_nnz   row pointer of C
//   allocate_C       C (Precise) / C-bar sized by the product bound (Upper)
//   compute          numeric merge, all tiers
//   copy             C-bar -> C compaction (Upper only)
//   cleanup          scratch release
// The host reads back only small counters (bin sizes, nnz(C), planner task
// counts, tile counts) between phases.
#include <stdio.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "brm_device.h"
#include "pipeline.h"

namespace brm {

Params& params() {
    static Params p{0, MED_CAP, 0, MED_CAP, 0, 0, GROUP_CFG_BITMAP, -1, 0, 1, 0, (int64_t)1 << 30, 0, 1, 1, 1, 1};
    return p;
}

namespace {
thread_local std::string t_err;
thread_local brm_stats t_stats;
thread_local int t_device = 0;
thread_local int64_t t_ws_budget = 0;
thread_local bool t_side_streams = true;
}  // namespace

brm_stats& thread_stats() { return t_stats; }
int64_t thread_ws_budget() { return t_ws_budget; }
void set_thread_ws_budget(int64_t bytes) { t_ws_budget = bytes; }
bool thread_side_streams() { return t_side_streams; }
void set_thread_side_streams(bool on) { t_side_streams = on; }
int thread_device() { return t_device; }
void set_thread_device(int device) { t_device = device; }

void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_err = buf;
}

int hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return BRM_OK;
    if (e == hipErrorOutOfMemory) {
        set_err("%s: out of device memory", what);
        return BRM_E_NOMEM;
    }
    set_err("%s: %s", what, hipGetErrorString(e));
    return BRM_E_DEVICE;
}

int use_device() {
    HIPC(hipSetDevice(t_device));
    return BRM_OK;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Phase events are recycled through a per-thread pool: creating and
// destroying some twenty events per call costs more host time than a small
// product's kernels.
namespace {
thread_local std::vector<hipEvent_t> t_event_pool;
hipEvent_t event_get() {
    if (!t_event_pool.empty()) {
        hipEvent_t e = t_event_pool.back();
        t_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
void event_put(hipEvent_t e) {
    if (!e) return;
    if (t_event_pool.size() < 256) t_event_pool.push_back(e);
    else (void)hipEventDestroy(e);
}
}  // namespace

PhaseClock::~PhaseClock() {
    for (auto& m : marks) {
        event_put(m.second.first);
        event_put(m.second.second);
    }
    event_put(open_ev);
}

void PhaseClock::begin(int p) {
    if (!on) return;
    end();
    open_ev = event_get();
    if (!open_ev) return;
    (void)hipEventRecord(open_ev, s);
    open_phase = p;
}

void PhaseClock::end() {
    if (open_phase < 0 || !open_ev) return;
    hipEvent_t e1 = event_get();
    if (e1) {
        (void)hipEventRecord(e1, s);
        marks.push_back({open_phase, {open_ev, e1}});
    } else {
        event_put(open_ev);
    }
    open_ev = nullptr;
    open_phase = -1;
}

void PhaseClock::collect(double* out) {
    end();
    for (int i = 0; i < P_N; ++i) out[i] = 0.0;
    for (auto& m : marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, m.second.first, m.second.second) == hipSuccess) out[m.first] += ms * 1e-3;
    }
}

// Small read-backs go through a pinned buffer of the calling thread (a copy
// into pageable memory is staged by the runtime and takes longer), several
// of them behind one synchronisation.  The buffer is kept until the process
// ends.
constexpr size_t D2H_PINNED = 4096;

int d2h_parts(std::initializer_list<D2HPart> parts, hipStream_t s) {
    static thread_local unsigned char* pinned = nullptr;
    size_t tot = 0;
    for (const D2HPart& q : parts) tot += (q.bytes + 15) & ~(size_t)15;
    if (tot <= D2H_PINNED && !pinned) {
        void* ptr = nullptr;
        if (hipHostMalloc(&ptr, D2H_PINNED, hipHostMallocDefault) == hipSuccess) pinned = (unsigned char*)ptr;
        else (void)hipGetLastError();
    }
    if (tot > D2H_PINNED || !pinned) {
        for (const D2HPart& q : parts) HIPC(hipMemcpyAsync(q.dst, q.src, q.bytes, hipMemcpyDeviceToHost, s));
        HIPC(hipStreamSynchronize(s));
        return BRM_OK;
    }
    size_t o = 0;
    for (const D2HPart& q : parts) {
        HIPC(hipMemcpyAsync(pinned + o, q.src, q.bytes, hipMemcpyDeviceToHost, s));
        o += (q.bytes + 15) & ~(size_t)15;
    }
    HIPC(hipStreamSynchronize(s));
    o = 0;
    for (const D2HPart& q : parts) {
        memcpy(q.dst, pinned + o, q.bytes);
        o += (q.bytes + 15) & ~(size_t)15;
    }
    return BRM_OK;
}

int d2h(void* dst, const void* src, size_t bytes, hipStream_t s) { return d2h_parts({{dst, src, bytes}}, s); }

int check_dev_err(const int* derr, hipStream_t s, const char* stage) {
    int e = 0;
    BRMC(d2h(&e, derr, sizeof(int), s));
    return dev_err_status(e, stage);
}

int dev_err_status(int e, const char* stage) {
    if (e & DERR_INVALID) {
        set_err("%s: malformed CSR input (row pointers, column range or unsorted/duplicate B columns)", stage);
        return BRM_E_INVALID;
    }
    if (e) {
        set_err("%s: internal consistency check failed (flags 0x%x)", stage, e);
        return BRM_E_INTERNAL;
    }
    return BRM_OK;
}

namespace {

// Host-side trace of the orchestration (BRM_TRACE=1 in the environment):
// elapsed milliseconds since the start of the call, after synchronising the
// stream, so gaps between kernels become visible.
struct HostTrace {
    bool on;
    double t0;
    hipStream_t s;
    explicit HostTrace(hipStream_t st) : on(getenv("BRM_TRACE") != nullptr), s(st) {
        t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    void mark(const char* what, int64_t x = -1) {
        if (!on) return;
        const double t1 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        (void)hipStreamSynchronize(s);
        const double t2 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        fprintf(stderr, "[brm] %9.3f ms (+%8.3f sync) %s %lld\n", (t1 - t0) * 1e3, (t2 - t1) * 1e3, what,
                (long long)x);
    }
};

// A second stream of the calling thread (default priority, created once per
// device and slot and kept until the process ends) for work that runs beside
// the main stream: fork() makes it wait for what the main stream has queued,
// mark() records where its work ends, join() makes the main stream wait for
// it.  When not `on`, everything runs on the main stream.  Slots are
// independent streams (0: the short and medium tiers, 1: alternate long-row
// waves).  The destructor waits for the side stream, so no early return
// leaves its kernels running over buffers about to be released.
class SideStream {
  public:
    SideStream(hipStream_t main, bool on, int slot = 0) : main_(main), on_(on && thread_side_streams()), slot_(slot) {}
    ~SideStream() {
        if (forked_) (void)hipStreamSynchronize(side_);
    }
    hipStream_t stream() const { return forked_ ? side_ : main_; }
    int fork() {
        if (!on_) return BRM_OK;
        int dev = 0;
        HIPC(hipGetDevice(&dev));
        static thread_local std::vector<std::pair<int, Res>> per_dev;  // (device * 2 + slot)
        Res* r = nullptr;
        for (auto& pd : per_dev)
            if (pd.first == dev * 2 + slot_) r = &pd.second;
        if (!r) {
            Res n{};
            HIPC(hipStreamCreateWithFlags(&n.side, hipStreamNonBlocking));
            HIPC(hipEventCreateWithFlags(&n.e0, hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&n.e1, hipEventDisableTiming));
            per_dev.push_back({dev * 2 + slot_, n});
            r = &per_dev.back().second;
        }
        side_ = r->side;
        e1_ = r->e1;
        HIPC(hipEventRecord(r->e0, main_));
        HIPC(hipStreamWaitEvent(side_, r->e0, 0));
        forked_ = true;
        return BRM_OK;
    }
    int mark() {
        if (forked_) HIPC(hipEventRecord(e1_, side_));
        return BRM_OK;
    }
    int join() {
        if (forked_) HIPC(hipStreamWaitEvent(main_, e1_, 0));
        return BRM_OK;
    }

  private:
    struct Res {
        hipStream_t side;
        hipEvent_t e0, e1;
    };
    hipStream_t main_;
    bool on_;
    int slot_;
    hipStream_t side_ = nullptr;
    hipEvent_t e1_ = nullptr;
    bool forked_ = false;
};

// ---------------------------------------------------------------- long rows
int run_long(const RowArgs& a, const int* rows_d, int nlong, const OutArgs& o, const int64_t* final_off,
             hipStream_t s, Scratch& sc) {
    if (nlong <= 0) return BRM_OK;
    HostTrace tr(s);
    tr.mark("run_long start", nlong);
    int64_t* ip_d;
    int* len_d;
    BRMC(sc.alloc(&ip_d, nlong, "long-row ip"));
    BRMC(sc.alloc(&len_d, nlong, "long-row len"));
    HIPC(launch_long_info(rows_d, nlong, a, ip_d, len_d, s));
    std::vector<int64_t> ip(nlong);
    std::vector<int> len(nlong), rows(nlong);
    BRMC(d2h(ip.data(), ip_d, sizeof(int64_t) * nlong, s));
    BRMC(d2h(len.data(), len_d, sizeof(int) * nlong, s));
    BRMC(d2h(rows.data(), rows_d, sizeof(int) * nlong, s));
    sc.free_now(ip_d);
    sc.free_now(len_d);

    // Wave partition by workspace budget: the ws_budget_bytes knob, else the
    // calling thread's share (set by the request server's workers), else 70%
    // of the memory this call can get, quantised to 4 GiB so that repeated
    // calls of one product plan identical waves.
    int64_t budget = params().ws_budget_bytes;
    if (budget <= 0) budget = thread_ws_budget();
    const int64_t arena = arena_remaining();
    if (arena >= 0) {
        // the arena bounds everything this function allocates: per long row,
        // 28 bytes of wave tables beside row_bytes(), and alignment slack
        budget = arena - (int64_t)28 * nlong - (int64_t)64 * 1024;
        if (budget < 0) {
            set_err("long rows need more scratch memory than the no-allocation entry's fixed buffer holds");
            return BRM_E_NOMEM;
        }
    } else if (budget <= 0) {
        size_t fr = 0, tot = 0;
        HIPC(hipMemGetInfo(&fr, &tot));
        const int64_t avail = (int64_t)(fr + dev_cached_bytes());
        budget = std::min(avail / 10 * 7, (int64_t)tot / 2);
        budget = budget / (int64_t(4) << 30) * (int64_t(4) << 30);
        if (budget < (int64_t)(1) << 30) budget = (int64_t)1 << 30;
        // Keep the previous call's budget while it still fits: the free
        // memory seen here moves a little from call to call, and a budget one
        // quantum off would re-partition the waves, resize every buffer and
        // defeat the allocator's cache.
        static thread_local int64_t last = 0;
        if (last > 0 && last <= avail / 10 * 8 && last <= (int64_t)tot / 2) budget = last;
        last = budget;
    }
    const int group_cfg = params().group_cfg;
    const bool bitmap = group_cfg == GROUP_CFG_BITMAP;
    const int pass_bits = bitmap ? BPASS_BITS : PASS_BITS;
    // Workspace elements row k stores in buffer b.  Pass j writes its nodes
    // below the root into buffer j & 1 and the root writes C, so with the
    // bitmap engine a row of one pass stores nothing and a row of two passes
    // only pass-1 nodes (buffer 1).  The merge-path engine's leaves may use
    // either buffer.
    // tree height of every row (ceil log2 of its list count), computed once
    std::vector<int> hk(nlong);
    for (int k = 0; k < nlong; ++k) hk[k] = len[k] <= 1 ? 0 : 32 - __builtin_clz((unsigned)(len[k] - 1));
    auto ws_need = [&](int k, int b) -> int64_t {
        if (!bitmap) return ip[k];
        const int H = hk[k];
        const int J = H == 0 ? 1 : (H + pass_bits - 1) / pass_bits;
        return (b == 1 ? J >= 2 : J >= 3) ? ip[k] : 0;
    };
    std::vector<int64_t> rbytes(nlong);
    for (int k = 0; k < nlong; ++k)
        rbytes[k] = 12 * (ws_need(k, 0) + ws_need(k, 1)) + 8 * ((int64_t)2 << hk[k]) +
                    (int64_t)(sizeof(GroupTask) + sizeof(LeafTask) + 2 * sizeof(Child)) * (len[k] + 16);

    tr.mark("long info read back");
    // ---- partition into waves and size every buffer for the largest wave
    struct Wave {
        int k0, k1;
        int64_t ws_tot[2], sl_tot, leaf_cap, child_cap;
        std::vector<int64_t> wsb, slb, seg;  // wsb: two per row (buffers 0 and 1)
    };
    std::vector<Wave> waves;
    int64_t mx_ws[2] = {0, 0}, mx_sl = 0, mx_leaf = 0, mx_child = 0, mx_task = 0;
    int mx_rows = 0;
  for (;;) {  // (the arena case re-plans with a smaller budget until the buffers fit)
    waves.clear();
    mx_ws[0] = mx_ws[1] = mx_sl = mx_leaf = mx_child = mx_task = 0;
    mx_rows = 0;
    for (int k0 = 0; k0 < nlong;) {
        int k1 = k0;
        int64_t used = 0;
        while (k1 < nlong && (k1 == k0 || used + rbytes[k1] <= budget)) used += rbytes[k1++];
        Wave w;
        w.k0 = k0;
        w.k1 = k1;
        w.ws_tot[0] = w.ws_tot[1] = w.sl_tot = w.leaf_cap = w.child_cap = 0;
        w.wsb.resize(2 * (size_t)(k1 - k0));
        w.slb.resize(k1 - k0);
        std::vector<int64_t> capP(MAX_PASSES + 1, 0);
        for (int k = 0; k < k1 - k0; ++k) {
            const int Lr = len[k0 + k];
            const int H = hk[k0 + k];
            const int J = H == 0 ? 1 : (H + pass_bits - 1) / pass_bits;
            if (J > MAX_PASSES) {
                set_err("long row with 2^%d lists exceeds the supported tree height", H);
                return BRM_E_INVALID;
            }
            w.wsb[2 * k] = w.ws_tot[0];
            w.wsb[2 * k + 1] = w.ws_tot[1];
            w.slb[k] = w.sl_tot;
            w.ws_tot[0] += ws_need(k0 + k, 0);
            w.ws_tot[1] += ws_need(k0 + k, 1);
            w.sl_tot += (int64_t)2 << H;
            w.leaf_cap += Lr;
            w.child_cap += 2 * (int64_t)Lr + 2;
            for (int j = 1; j <= J; ++j) capP[j] += (int64_t)(Lr >> std::min(j * pass_bits, H)) + 2;
        }
        w.seg.assign(MAX_PASSES + 2, 0);
        for (int j = 0; j <= MAX_PASSES; ++j) w.seg[j + 1] = w.seg[j] + capP[j];
        mx_ws[0] = std::max(mx_ws[0], w.ws_tot[0]);
        mx_ws[1] = std::max(mx_ws[1], w.ws_tot[1]);
        mx_sl = std::max(mx_sl, w.sl_tot);
        mx_leaf = std::max(mx_leaf, w.leaf_cap);
        mx_child = std::max(mx_child, w.child_cap);
        mx_task = std::max(mx_task, w.seg[MAX_PASSES + 1]);
        mx_rows = std::max(mx_rows, k1 - k0);
        waves.push_back(std::move(w));
        k0 = k1;
    }
    if (arena < 0) break;
    // bytes of the buffers allocated below (each sized for its largest wave)
    const int64_t need = 12 * (mx_ws[0] + mx_ws[1]) + 8 * mx_sl + (int64_t)sizeof(LeafTask) * mx_leaf +
                         (int64_t)sizeof(Child) * mx_child + (int64_t)sizeof(GroupTask) * mx_task +
                         28 * (int64_t)mx_rows + 8 * (CNT_N + MAX_PASSES + 2) + 16 * 256;
    if (need <= arena) break;
    if ((int)waves.size() == nlong) {
        set_err("a long row needs more scratch memory than the no-allocation entry's fixed buffer holds");
        return BRM_E_NOMEM;
    }
    budget = budget / 4 * 3;
  }

    // Buffers are rounded up to one eighth of their next power of two (at
    // most 12.5% more than needed, with a small floor), so that products of
    // similar size reuse each other's cached blocks.
    auto round_geo = [](int64_t n, int64_t floor) {
        if (n <= floor) return floor;
        int64_t p2 = 1;
        while (p2 < n) p2 <<= 1;
        const int64_t g = p2 / 8;
        return (n + g - 1) / g * g;
    };
    if (arena < 0) {
        mx_ws[0] = round_geo(mx_ws[0], (int64_t)1 << 20);
        mx_ws[1] = round_geo(mx_ws[1], (int64_t)1 << 20);
        mx_sl = round_geo(mx_sl, (int64_t)1 << 16);
        mx_leaf = round_geo(mx_leaf, (int64_t)1 << 14);
        mx_child = round_geo(mx_child, (int64_t)1 << 14);
        mx_task = round_geo(mx_task, (int64_t)1 << 14);
        mx_rows = (int)round_geo(mx_rows, (int64_t)1 << 10);
    }
    Scratch wsc;
    int *wcol0, *wcol1;
    double *wval0, *wval1;
    LeafTask* leaves;
    BRMC(wsc.alloc(&wcol0, mx_ws[0], "workspace"));
    BRMC(wsc.alloc(&wcol1, mx_ws[1], "workspace"));
    BRMC(wsc.alloc(&wval0, mx_ws[0], "workspace"));
    BRMC(wsc.alloc(&wval1, mx_ws[1], "workspace"));
    BRMC(wsc.alloc(&leaves, mx_leaf, "leaf tasks"));
    // Per-wave tables.  Bitmap engine: each pass's tasks run largest node
    // first, so that the longest merges start early and the short ones fill
    // the end of the pass (not in the fixed-buffer arena, whose budget does
    // not hold the sort's buffers).
    struct WaveTables {
        int* rows_w = nullptr;
        int64_t *wsb_d = nullptr, *slb_d = nullptr, *seg_d = nullptr, *slots = nullptr, *counters = nullptr;
        GroupTask* tasks = nullptr;
        Child* children = nullptr;
        uint32_t *tkey = nullptr, *tkey_out = nullptr;
        int *tidx_in = nullptr, *tidx_out = nullptr;
        unsigned char* ttmp = nullptr;
    };
    const size_t ttmp_bytes = bitmap && arena < 0 && mx_task <= INT32_MAX ? order_tasks_tmp_bytes((int)mx_task) : 0;
    // Consecutive waves overlap on two streams (bitmap engine with ordered
    // tasks, more than one wave): each parity has its own tables, and the
    // workspace buffers are handed over by events -- wave w+1's pass 1
    // (writes buffer 1) waits for the last pass of wave w that touches buffer
    // 1, its pass 2 (writes buffer 0) for the end of wave w.  Wave w's last
    // pass, the roots of its heaviest rows, then runs beside wave w+1's
    // planning and first pass.
    const bool overlap = bitmap && arena < 0 && ttmp_bytes > 0 && waves.size() > 1 && thread_side_streams();
    WaveTables wt[2];
    for (int par = 0; par < (overlap ? 2 : 1); ++par) {
        WaveTables& t = wt[par];
        BRMC(wsc.alloc(&t.rows_w, mx_rows, "wave rows"));
        BRMC(wsc.alloc(&t.wsb_d, 2 * (int64_t)mx_rows, "wave ws_base"));
        BRMC(wsc.alloc(&t.slb_d, mx_rows, "wave slot_base"));
        BRMC(wsc.alloc(&t.seg_d, MAX_PASSES + 2, "task segments"));
        BRMC(wsc.alloc(&t.slots, mx_sl, "length slots"));
        BRMC(wsc.alloc(&t.counters, CNT_N, "planner counters"));
        BRMC(wsc.alloc(&t.tasks, mx_task, "group tasks"));
        BRMC(wsc.alloc(&t.children, mx_child, "group children"));
        if (ttmp_bytes > 0) {
            BRMC(wsc.alloc(&t.tkey, mx_task, "task keys"));
            BRMC(wsc.alloc(&t.tkey_out, mx_task, "task keys"));
            BRMC(wsc.alloc(&t.tidx_in, mx_task, "task order"));
            BRMC(wsc.alloc(&t.tidx_out, mx_task, "task order"));
            BRMC(wsc.alloc(&t.ttmp, (int64_t)ttmp_bytes, "task sort"));
        }
    }
    tr.mark("workspace allocated", mx_ws[0] + mx_ws[1]);
    int* const wcol[2] = {wcol0, wcol1};
    double* const wval[2] = {wval0, wval1};
    int* derr = o.err;
    const LdsShape sh = lds_shape(params().lds_cfg);
    SideStream wave2(s, overlap, 1);
    BRMC(wave2.fork());
    struct Hand {
        hipEvent_t b1_free = nullptr, done = nullptr;  // handed to the next wave
    };
    static thread_local std::vector<std::pair<int, Hand>> hand_pool;  // per device, two per parity
    Hand hand[2];
    if (overlap) {
        int dev = 0;
        HIPC(hipGetDevice(&dev));
        for (int par = 0; par < 2; ++par) {
            Hand* h = nullptr;
            for (auto& hp : hand_pool)
                if (hp.first == dev * 2 + par) h = &hp.second;
            if (!h) {
                Hand n;
                HIPC(hipEventCreateWithFlags(&n.b1_free, hipEventDisableTiming));
                HIPC(hipEventCreateWithFlags(&n.done, hipEventDisableTiming));
                hand_pool.push_back({dev * 2 + par, n});
                h = &hand_pool.back().second;
            }
            hand[par] = *h;
        }
    }
    // the work counters of a wave, read once its stream has finished it
    struct PendingStats {
        bool on = false;
        int par = 0;
        hipStream_t st = nullptr;
        std::vector<int64_t> cnt;
        int64_t ntask = 0;
        int passes = 0;
    };
    auto read_stats = [&](PendingStats& ps) -> int {
        if (!ps.on) return BRM_OK;
        ps.on = false;
        std::vector<int64_t> c2(CNT_N);
        BRMC(d2h(c2.data(), wt[ps.par].counters, sizeof(int64_t) * CNT_N, ps.st));
        t_stats.long_waves += 1;
        t_stats.leaf_tasks += ps.cnt[CNT_LEAF];
        t_stats.leaf_ip += ps.cnt[CNT_LEAF_IP];
        t_stats.leaf_out += c2[CNT_LEAF_OUT];
        t_stats.group_tasks += ps.ntask;
        t_stats.group_rounds += c2[CNT_ROUNDS];
        t_stats.group_in += c2[CNT_GRP_IN];
        t_stats.group_out += c2[CNT_GRP_OUT];
        t_stats.passes = std::max<int64_t>(t_stats.passes, ps.passes);
        if (tr.on && bitmap) {
            fprintf(stderr, "[brm] bitmap engine cycles by phase (sum over workgroups):");
            for (int p = 0; p < N_PHASES; ++p) fprintf(stderr, " %.3g", (double)c2[CNT_PH0 + p]);
            fprintf(stderr, "\n");
        }
        return BRM_OK;
    };
    PendingStats prev_stats;

    for (size_t wi = 0; wi < waves.size(); ++wi) {
        const Wave& w = waves[wi];
        const int par = overlap ? (int)(wi & 1) : 0;
        const hipStream_t ws = par ? wave2.stream() : s;
        WaveTables& t = wt[par];
        const int nw = w.k1 - w.k0;
        HIPC(hipMemcpyAsync(t.rows_w, rows.data() + w.k0, sizeof(int) * nw, hipMemcpyHostToDevice, ws));
        HIPC(hipMemcpyAsync(t.wsb_d, w.wsb.data(), sizeof(int64_t) * 2 * nw, hipMemcpyHostToDevice, ws));
        HIPC(hipMemcpyAsync(t.slb_d, w.slb.data(), sizeof(int64_t) * nw, hipMemcpyHostToDevice, ws));
        HIPC(hipMemcpyAsync(t.seg_d, w.seg.data(), sizeof(int64_t) * (MAX_PASSES + 2), hipMemcpyHostToDevice, ws));
        HIPC(launch_fill_i64(t.counters, CNT_N, 0, ws));
        HIPC(launch_fill_i64(t.slots, w.sl_tot, 0, ws));

        LongArgs la;
        la.rows = t.rows_w;
        la.nrows = nw;
        la.ws_base = t.wsb_d;
        la.slot_base = t.slb_d;
        la.arpt = a.arpt;
        la.acol = a.acol;
        la.aval = a.aval;
        la.brpt = a.brpt;
        la.ipA = a.ipA;
        la.final_off = final_off;
        la.counters = t.counters;
        la.leaves = leaves;
        la.leaves_cap = w.leaf_cap;
        la.tasks = t.tasks;
        la.task_key = t.tkey;
        la.task_seg = t.seg_d;
        la.children = t.children;
        la.child_cap = w.child_cap;
        la.frontier_cap = bitmap ? 0 : std::min(params().leaf_cap, sh.cap);
        la.lmax = sh.lmax;
        la.log_lmax = sh.log_lmax;
        la.pass_bits = pass_bits;
        tr.mark("wave uploaded", nw);
        HIPC(launch_plan(la, derr, ws));
        tr.mark("planned");
        std::vector<int64_t> cnt(CNT_N);
        BRMC(d2h(cnt.data(), t.counters, sizeof(int64_t) * CNT_N, ws));
        BRMC(check_dev_err(derr, ws, "long-row planner"));

        HIPC(launch_leaf(leaves, cnt[CNT_LEAF], a, wcol, wval, t.slots, t.counters, o, ws));
        tr.mark("leaves", cnt[CNT_LEAF]);
        const int passes = (int)cnt[CNT_MAXP];
        // the last pass that reads or writes buffer 1: passes j writing
        // nodes below the root into buffer j & 1, reading buffer (j-1) & 1
        int b1_last = 0;
        for (int j = 1; j <= passes && j <= MAX_PASSES; ++j)
            if ((j % 2 == 0) || j < passes) b1_last = j;
        if (overlap && b1_last == 0) {
            // no pass of this wave uses buffer 1: it is free once the
            // previous wave has let it go
            if (wi > 0) HIPC(hipStreamWaitEvent(ws, hand[par ^ 1].b1_free, 0));
            HIPC(hipEventRecord(hand[par].b1_free, ws));
        }
        int64_t ntask = 0;
        for (int j = 1; j <= passes && j <= MAX_PASSES; ++j) {
            if (overlap && wi > 0 && j <= 2) HIPC(hipStreamWaitEvent(ws, j == 1 ? hand[par ^ 1].b1_free : hand[par ^ 1].done, 0));
            const int* order = nullptr;
            if (t.tkey && cnt[CNT_TASK0 + j] > 1) {
                HIPC(order_tasks(t.tkey + w.seg[j], (int)cnt[CNT_TASK0 + j], t.tkey_out, t.tidx_in, t.tidx_out, t.ttmp,
                                 ttmp_bytes, ws));
                order = t.tidx_out;
            }
            HIPC(launch_group(t.tasks + w.seg[j], cnt[CNT_TASK0 + j], t.children, a, wcol, wval, t.slots, t.counters, o,
                              ws, group_cfg, order));
            ntask += cnt[CNT_TASK0 + j];
            tr.mark("group pass", cnt[CNT_TASK0 + j]);
            if (overlap && j == b1_last) HIPC(hipEventRecord(hand[par].b1_free, ws));
        }
        if (overlap) {
            // a wave of one pass still waits for the previous wave's end
            if (wi > 0 && passes < 2) HIPC(hipStreamWaitEvent(ws, hand[par ^ 1].done, 0));
            HIPC(hipEventRecord(hand[par].done, ws));
        }
        // the previous wave's counters (its stream finishes it beside this
        // wave's passes), or this wave's when waves do not overlap
        BRMC(read_stats(prev_stats));
        prev_stats.on = true;
        prev_stats.par = par;
        prev_stats.st = ws;
        prev_stats.cnt = cnt;
        prev_stats.ntask = ntask;
        prev_stats.passes = passes;
        if (!overlap) BRMC(read_stats(prev_stats));
        dev_flush_deferred();
    }
    BRMC(read_stats(prev_stats));
    BRMC(wave2.mark());
    BRMC(wave2.join());
    return BRM_OK;
}

}  // namespace

// ---------------------------------------------------------------- pipeline
namespace {
// Row bins (device) under the given tier parameters; the sizes stay on the device.
hipError_t bin_launch(Pattern& p, int lds_cfg, int med_cap, int wmed_cap, int force_long, hipStream_t s) {
    const LdsShape shape = lds_shape(lds_cfg);
    p.lds_cfg = lds_cfg;
    p.med_cap = std::min(med_cap, shape.cap);
    if (wmed_cap < 0) wmed_cap = params().wmed_engine == 1 ? rmed_auto_cap(p.a.N) : 512;
    // rows above WMED_CAP only where the register tier ranks with a bitmap
    p.wmed_cap = std::min(wmed_cap, params().wmed_engine == 1 ? rmed_max_cap(p.a.N) : WMED_CAP);
    hipError_t e = launch_fill_i32(p.bin_cnt, BIN_COUNT, 0, s);
    if (e != hipSuccess) return e;
    return launch_rowstats(p.a, p.med_cap, shape.lmax, p.wmed_cap, force_long, params().hash_cap, p.iprow, p.rowcnt,
                           p.bins, p.bin_cnt, p.derr, s);
}
}  // namespace

int pattern_prepare(const DevIn& in, hipStream_t s, PhaseClock* clk, Pattern& p) {
    if (in.M < 0 || in.K < 0 || in.N < 0 || in.nnzA < 0 || in.nnzB < 0) {
        set_err("negative dimension");
        return BRM_E_INVALID;
    }
    if (in.nnzA > INT32_MAX || in.nnzB > INT32_MAX) {
        set_err("nnz(A) and nnz(B) must fit int32 row pointers");
        return BRM_E_INVALID;
    }
    const int M = in.M;
    if (clk) clk->begin(P_PRE);
    Scratch& sc = p.sc;
    BRMC(sc.alloc(&p.derr, 1, "error flag"));
    BRMC(sc.alloc(&p.bin_cnt, BIN_COUNT, "bin counters"));
    BRMC(sc.alloc(&p.ipA, in.nnzA + 1, "ipA"));
    BRMC(sc.alloc(&p.iprow, M, "row ip"));
    BRMC(sc.alloc(&p.rowcnt, M, "row nnz"));
    if (p.crpt_ext) p.crpt = p.crpt_ext;
    else BRMC(sc.alloc(&p.crpt, (int64_t)M + 1, "crpt"));
    BRMC(sc.alloc(&p.bins, (BIN_COUNT - 1) * (int64_t)M, "row bins"));
    HIPC(launch_fill_i32(p.derr, 1, 0, s));
    p.nnzA = in.nnzA;
    p.nnzB = in.nnzB;
    RowArgs& a = p.a;
    a.arpt = in.arpt;
    a.acol = in.acol;
    a.aval = in.aval;
    a.brpt = in.brpt;
    a.bcol = in.bcol;
    a.bval = in.bval;
    a.ipA = p.ipA;
    a.M = in.M;
    a.K = in.K;
    a.N = in.N;
    HIPC(launch_validate(a, in.nnzA, in.b_checked ? -1 : in.nnzB, p.derr, s));
    if (!in.b_checked) HIPC(launch_validate_sorted(in.brpt, in.bcol, in.K, in.nnzB, p.derr, s));

    // The product scan reads no B row pointer outside the array and the
    // row binning skips every row once validation has flagged the input, so
    // one host round trip after binning serves validation, the product
    // total and the bin sizes.
    if (clk) clk->begin(P_FLOP);
    HIPC(scan_alen(in.acol, in.brpt, in.K, in.nnzA, p.ipA, s));

    if (clk) clk->begin(P_LB);
    const LdsShape shape = lds_shape(params().lds_cfg);
    HIPC(bin_launch(p, params().lds_cfg, std::min(params().med_cap, shape.cap), params().wmed_cap,
                    params().force_long, s));
    int e = 0;
    BRMC(d2h_parts({{&e, p.derr, sizeof(int)}, {&p.total_ip, p.ipA + in.nnzA, sizeof(int64_t)},
                    {p.bc, p.bin_cnt, sizeof(p.bc)}},
                   s));
    return dev_err_status(e, "validation");
}

int pattern_bin(Pattern& p, int lds_cfg, int med_cap, int wmed_cap, int force_long, hipStream_t s) {
    HIPC(bin_launch(p, lds_cfg, med_cap, wmed_cap, force_long, s));
    BRMC(d2h(p.bc, p.bin_cnt, sizeof(p.bc), s));
    return BRM_OK;
}

int pattern_symbolic(Pattern& p, hipStream_t s, PhaseClock* clk, bool stash_short, bool read_nnz) {
    if (clk) clk->begin(P_SYM);
    // Long rows, largest first (the heaviest rows' counts start early; the
    // long-row engine later takes them in this order too), counted on a side
    // stream beside the other tiers' counts.
    const int nlong = p.bc[BIN_LONG];
    const bool own_mem = arena_remaining() < 0;
    SideStream side(s, nlong > 0 && own_mem);
    BRMC(side.fork());
    if (nlong > 1 && own_mem) {
        const size_t tb = order_rows_tmp_bytes(nlong);
        if (tb > 0) {
            Scratch tsc;
            uint64_t *k0, *k1;
            int* ro;
            unsigned char* tmp;
            BRMC(tsc.alloc(&k0, nlong, "long-row keys"));
            BRMC(tsc.alloc(&k1, nlong, "long-row keys"));
            BRMC(tsc.alloc(&ro, nlong, "long-row order"));
            BRMC(tsc.alloc(&tmp, (int64_t)tb, "long-row sort"));
            HIPC(order_rows(const_cast<int*>(p.rows(BIN_LONG)), nlong, p.a.arpt, p.a.ipA, k0, k1, ro, tmp, tb,
                            side.stream()));
            // the queued sort still uses these: released when the call ends,
            // after its final synchronisation
            for (void* q : {(void*)k0, (void*)k1, (void*)ro, (void*)tmp}) {
                tsc.release(q);
                dev_defer_free(q);
            }
        }
    }
    HIPC(launch_long_count(p.rows(BIN_LONG), nlong, p.a, p.rowcnt, side.stream()));
    BRMC(side.mark());
    const int64_t n32 = (int64_t)p.bc[BIN_TINY] * 32, n64 = (int64_t)p.bc[BIN_SHORT] * 64;
    p.stashed = stash_short && n32 + n64 > 0 && 12 * (n32 + n64) <= SHORT_STASH_MAX_BYTES;
    if (p.stashed) {
        BRMC(p.sc.alloc(&p.stash_col[0], n32, "short-row stash"));
        BRMC(p.sc.alloc(&p.stash_val[0], n32, "short-row stash"));
        BRMC(p.sc.alloc(&p.stash_col[1], n64, "short-row stash"));
        BRMC(p.sc.alloc(&p.stash_val[1], n64, "short-row stash"));
        OutArgs o{};
        o.err = p.derr;
        o.upper = 1;
        o.rowlen = p.rowcnt;
        o.slot = 32;
        o.ccol = p.stash_col[0];
        o.cval = p.stash_val[0];
        HIPC(launch_short_numeric(p.rows(BIN_TINY), p.bc[BIN_TINY], p.a, o, s, 32));
        o.slot = 64;
        o.ccol = p.stash_col[1];
        o.cval = p.stash_val[1];
        HIPC(launch_short_numeric(p.rows(BIN_SHORT), p.bc[BIN_SHORT], p.a, o, s));
    } else {
        HIPC(launch_short_count(p.rows(BIN_TINY), p.bc[BIN_TINY], p.a, p.rowcnt, s, 32));
        HIPC(launch_short_count(p.rows(BIN_SHORT), p.bc[BIN_SHORT], p.a, p.rowcnt, s));
    }
    HIPC(launch_wmed_count(p.rows(BIN_WMED), p.bc[BIN_WMED], p.a, p.rowcnt, s, std::min(p.wmed_cap, WMED_SPLIT0)));
    HIPC(launch_wmed_count(p.rows(BIN_WMED2), p.bc[BIN_WMED2], p.a, p.rowcnt, s, std::min(p.wmed_cap, WMED_SPLIT)));
    HIPC(launch_wmed_count(p.rows(BIN_WMED3), p.bc[BIN_WMED3], p.a, p.rowcnt, s, std::min(p.wmed_cap, WMED_SPLIT2)));
    HIPC(launch_wmed_count(p.rows(BIN_WMED4), p.bc[BIN_WMED4], p.a, p.rowcnt, s, p.wmed_cap));
    HIPC(launch_medium_count(p.rows(BIN_MEDIUM), p.bc[BIN_MEDIUM], p.a, p.rowcnt, s));
    HIPC(launch_hash_count(p.rows(BIN_HASH), p.bc[BIN_HASH], p.a, p.rowcnt, s, HASH_SMALL));
    HIPC(launch_hash_count(p.rows(BIN_HASH2), p.bc[BIN_HASH2], p.a, p.rowcnt, s, HASH_CAP));
    BRMC(side.join());
    if (clk) clk->begin(P_PNNZ);
    HIPC(scan_i64(p.rowcnt, p.a.M, p.crpt, s));
    p.nnzC = -1;
    if (read_nnz) BRMC(d2h(&p.nnzC, p.crpt + p.a.M, sizeof(int64_t), s));
    return BRM_OK;
}

namespace {
OutArgs precise_out(const Pattern& p, int* ccol, double* cval) {
    OutArgs o{};
    o.err = p.derr;
    o.upper = 0;
    o.crpt = p.crpt;
    o.ccol = ccol;
    o.cval = cval;
    return o;
}
}  // namespace

int pattern_numeric_tiers(const Pattern& p, int* ccol, double* cval, hipStream_t s) {
    const OutArgs o = precise_out(p, ccol, cval);
    if (p.stashed) {
        HIPC(launch_short_copy(p.rows(BIN_TINY), p.bc[BIN_TINY], p.stash_col[0], p.stash_val[0], p.crpt, ccol, cval,
                               s, 32));
        HIPC(launch_short_copy(p.rows(BIN_SHORT), p.bc[BIN_SHORT], p.stash_col[1], p.stash_val[1], p.crpt, ccol,
                               cval, s, 64));
    } else {
        HIPC(launch_short_numeric(p.rows(BIN_TINY), p.bc[BIN_TINY], p.a, o, s, 32));
        HIPC(launch_short_numeric(p.rows(BIN_SHORT), p.bc[BIN_SHORT], p.a, o, s));
    }
    HIPC(launch_wmed_numeric(p.rows(BIN_WMED), p.bc[BIN_WMED], p.a, o, s, std::min(p.wmed_cap, WMED_SPLIT0)));
    HIPC(launch_wmed_numeric(p.rows(BIN_WMED2), p.bc[BIN_WMED2], p.a, o, s, std::min(p.wmed_cap, WMED_SPLIT)));
    HIPC(launch_wmed_numeric(p.rows(BIN_WMED3), p.bc[BIN_WMED3], p.a, o, s, std::min(p.wmed_cap, WMED_SPLIT2)));
    HIPC(launch_wmed_numeric(p.rows(BIN_WMED4), p.bc[BIN_WMED4], p.a, o, s, p.wmed_cap));
    HIPC(launch_medium_numeric(p.rows(BIN_MEDIUM), p.bc[BIN_MEDIUM], p.a, o, s, p.lds_cfg));
    HIPC(launch_hash_numeric(p.rows(BIN_HASH), p.bc[BIN_HASH], p.a, o, s, HASH_SMALL));
    HIPC(launch_hash_numeric(p.rows(BIN_HASH2), p.bc[BIN_HASH2], p.a, o, s, HASH_CAP));
    return BRM_OK;
}

int pattern_numeric_long(const Pattern& p, int* ccol, double* cval, hipStream_t s) {
    if (p.bc[BIN_LONG] == 0) return BRM_OK;
    const OutArgs o = precise_out(p, ccol, cval);
    Scratch sc;
    return run_long(p.a, p.rows(BIN_LONG), p.bc[BIN_LONG], o, p.crpt, s, sc);
}

namespace {
// C-bar bytes one Upper row wave may use: the cbar_budget_bytes knob, or a
// quarter of the device memory this call can get (free plus the allocator's
// cached blocks), at most an eighth of the device, in whole GiB.
int upper_wave_budget(int64_t* out) {
    int64_t budget = params().cbar_budget_bytes;
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        HIPC(hipMemGetInfo(&fr, &tot));
        budget = std::min((int64_t)(fr + dev_cached_bytes()) / 4, (int64_t)tot / 8);
        budget = std::max(budget / (int64_t(1) << 30), (int64_t)1) * (int64_t(1) << 30);
    }
    *out = budget;
    return BRM_OK;
}

// BRMerge-Upper numeric phase and assembly (spgemm.h:293-445): every row is
// merged into C-bar at its product offset ipA[arpt[row]] (the reference's
// scan_floprC[row], :399-400) with its merged length in rowcnt, then C's row
// pointer is scanned (:413) and C-bar compacted into C (:427-445).
//
// The reference holds C-bar for every row at once (12 bytes per product,
// 251 GB for R-MAT scale 20).  Here rows run in waves of consecutive rows
// whose C-bar fits upper_wave_budget(): each wave is re-binned, merged into
// one reused C-bar buffer (addressed relative to the wave's first product),
// and compacted into a chunk of exactly its own nnz.  Once every wave is done
// C is allocated at nnz(C) and the chunks are copied into place.  A product
// whose C-bar fits one wave is compacted straight into C.
int run_upper(Pattern& p, int64_t max_nnz, hipStream_t s, PhaseClock& clk, int** ccol_out, double** cval_out,
              int64_t* nnz_out) {
    const int M = p.a.M;
    Scratch& sc = p.sc;
    clk.begin(P_ALLOC);
    int64_t budget = 0;
    BRMC(upper_wave_budget(&budget));
    const int64_t cap_ip = std::max<int64_t>(1, budget / 12);
    // Wave boundaries (row indices) and each wave's first product.
    std::vector<int> wb{0};
    std::vector<int64_t> wbase{0};
    int64_t wave_ip_max = p.total_ip;
    if (p.total_ip > cap_ip) {
        std::vector<int64_t> iprow(M);
        BRMC(d2h(iprow.data(), p.iprow, sizeof(int64_t) * M, s));
        int64_t acc = 0, base = 0;
        wave_ip_max = 0;
        for (int r = 0; r < M; ++r) {
            if (acc > 0 && iprow[r] > 0 && acc + iprow[r] > cap_ip) {  // empty rows never open a wave
                wb.push_back(r);
                base += acc;
                wbase.push_back(base);
                wave_ip_max = std::max(wave_ip_max, acc);
                acc = 0;
            }
            acc += iprow[r];
        }
        wave_ip_max = std::max(wave_ip_max, acc);
    }
    wb.push_back(M);
    const int nwaves = (int)wb.size() - 1;
    int* bcol_bar;
    double* bval_bar;
    BRMC(sc.alloc(&bcol_bar, wave_ip_max, "C-bar col"));
    BRMC(sc.alloc(&bval_bar, wave_ip_max, "C-bar val"));
    thread_stats().upper_waves = nwaves;

    int64_t* crpt_w = nullptr;
    if (nwaves > 1) {
        int mrows = 0;
        for (int w = 0; w < nwaves; ++w) mrows = std::max(mrows, wb[w + 1] - wb[w]);
        BRMC(sc.alloc(&crpt_w, (int64_t)mrows + 1, "wave row pointer"));
    }
    // Several waves: C grows wave by wave in reserved address ranges (physical
    // memory mapped as it grows, nothing moved); where the driver cannot
    // reserve them, each wave is compacted into an exact-size chunk and the
    // chunks are copied into C at the end.
    int* vcol = nullptr;
    double* vval = nullptr;
    int64_t voff = 0;
    if (nwaves > 1 && params().upper_vm) {
        vcol = (int*)vm_reserve(sizeof(int) * (size_t)p.total_ip, thread_device());
        if (vcol) sc.bufs.push_back(vcol);
        vval = vcol ? (double*)vm_reserve(sizeof(double) * (size_t)p.total_ip, thread_device()) : nullptr;
        if (vval) sc.bufs.push_back(vval);
        // map a first step of each now, so that a driver without physical
        // allocations for reserved ranges falls back before any wave runs
        if (vcol && (!vval || !vm_grow(vcol, 1) || !vm_grow(vval, 1))) {
            sc.free_now(vcol);
            if (vval) sc.free_now(vval);
            vcol = nullptr;
            vval = nullptr;
        }
    }
    // The long-row workspace of every wave gets the same budget, so that the
    // waves' buffers are sized alike and reused from the cache: twice C-bar's
    // (a wave's long rows store up to 24 bytes per product in the two
    // workspace buffers; with C-bar's budget alone they split into several
    // workspace waves), at most a quarter of the device.
    int64_t ws_budget = 0;
    if (nwaves > 1) {
        size_t fr = 0, tot = 0;
        HIPC(hipMemGetInfo(&fr, &tot));
        ws_budget = std::min(2 * budget, (int64_t)tot / 4);
    }
    struct WsBudgetScope {
        int64_t saved;
        explicit WsBudgetScope(int64_t b) : saved(thread_ws_budget()) {
            if (b > 0 && saved <= 0 && params().ws_budget_bytes <= 0) set_thread_ws_budget(b);
        }
        ~WsBudgetScope() { set_thread_ws_budget(saved); }
    } ws_scope(ws_budget);
    Scratch chunks;
    std::vector<int*> ch_col;
    std::vector<double*> ch_val;
    std::vector<int64_t> ch_n;
    const LdsShape shape = lds_shape(p.lds_cfg);
    for (int w = 0; w < nwaves; ++w) {
        RowArgs aw = p.a;
        aw.arpt = p.a.arpt + wb[w];
        aw.M = wb[w + 1] - wb[w];
        int bc[BIN_COUNT];
        const int* bins = p.bins;
        if (nwaves == 1) {
            std::copy(p.bc, p.bc + BIN_COUNT, bc);
        } else {
            // bins of this wave's rows (row ids relative to its first row)
            clk.begin(P_LB);
            HIPC(launch_fill_i32(p.bin_cnt, BIN_COUNT, 0, s));
            HIPC(launch_rowstats(aw, p.med_cap, shape.lmax, p.wmed_cap, params().force_long, params().hash_cap,
                                 p.iprow + wb[w],
                                 p.rowcnt + wb[w], p.bins, p.bin_cnt, p.derr, s));
            BRMC(d2h(bc, p.bin_cnt, sizeof(bc), s));
        }
        auto rows = [&](int bin) { return bins + (int64_t)(bin - 1) * aw.M; };
        clk.begin(P_COMPUTE);
        // C-bar holds products [wbase[w], wbase[w] + wave ip): kernels index it
        // by the global ipA, so its base pointers are shifted back by wbase[w].
        OutArgs o{};
        o.err = p.derr;
        o.upper = 1;
        o.ccol = bcol_bar - wbase[w];
        o.cval = bval_bar - wbase[w];
        o.rowlen = p.rowcnt + wb[w];
        // the short and medium tiers beside the long-row engine, as for Precise
        SideStream side(s, bc[BIN_LONG] > 0 && arena_remaining() < 0);
        BRMC(side.fork());
        const hipStream_t ts = side.stream();
        HIPC(launch_short_numeric(rows(BIN_TINY), bc[BIN_TINY], aw, o, ts, 32));
        HIPC(launch_short_numeric(rows(BIN_SHORT), bc[BIN_SHORT], aw, o, ts));
        HIPC(launch_wmed_numeric(rows(BIN_WMED), bc[BIN_WMED], aw, o, ts, std::min(p.wmed_cap, WMED_SPLIT0)));
        HIPC(launch_wmed_numeric(rows(BIN_WMED2), bc[BIN_WMED2], aw, o, ts, std::min(p.wmed_cap, WMED_SPLIT)));
        HIPC(launch_wmed_numeric(rows(BIN_WMED3), bc[BIN_WMED3], aw, o, ts, std::min(p.wmed_cap, WMED_SPLIT2)));
        HIPC(launch_wmed_numeric(rows(BIN_WMED4), bc[BIN_WMED4], aw, o, ts, p.wmed_cap));
        HIPC(launch_medium_numeric(rows(BIN_MEDIUM), bc[BIN_MEDIUM], aw, o, ts, p.lds_cfg));
        HIPC(launch_hash_numeric(rows(BIN_HASH), bc[BIN_HASH], aw, o, ts, HASH_SMALL));
        HIPC(launch_hash_numeric(rows(BIN_HASH2), bc[BIN_HASH2], aw, o, ts, HASH_CAP));
        BRMC(side.mark());
        BRMC(run_long(aw, rows(BIN_LONG), bc[BIN_LONG], o, nullptr, s, sc));
        BRMC(side.join());
        BRMC(check_dev_err(p.derr, s, "numeric"));
        if (nwaves == 1) {
            clk.begin(P_PNNZ);
            HIPC(scan_i64(p.rowcnt, M, p.crpt, s));
            int64_t nnzC = 0;
            BRMC(d2h(&nnzC, p.crpt + M, sizeof(int64_t), s));
            if (nnzC > max_nnz) {
                set_err("nnz(C) = %lld exceeds the int32 row pointer", (long long)nnzC);
                return BRM_E_OVERFLOW;
            }
            clk.begin(P_ALLOC);
            BRMC(sc.alloc(ccol_out, nnzC, "C col"));
            BRMC(sc.alloc(cval_out, nnzC, "C val"));
            clk.begin(P_COPY);
            HIPC(launch_compact(p.crpt, M, aw, o.ccol, o.cval, *ccol_out, *cval_out, nnzC, s));
            *nnz_out = nnzC;
            return BRM_OK;
        }
        clk.begin(P_PNNZ);
        HIPC(scan_i64(p.rowcnt + wb[w], aw.M, crpt_w, s));
        int64_t nw = 0;
        BRMC(d2h(&nw, crpt_w + aw.M, sizeof(int64_t), s));
        clk.begin(P_ALLOC);
        int* cc;
        double* cv;
        if (vcol) {
            if (!vm_grow(vcol, sizeof(int) * (size_t)(voff + nw)) ||
                !vm_grow(vval, sizeof(double) * (size_t)(voff + nw))) {
                set_err("mapping %lld more entries of C failed", (long long)nw);
                return BRM_E_NOMEM;
            }
            cc = vcol + voff;
            cv = vval + voff;
            voff += nw;
        } else {
            BRMC(chunks.alloc(&cc, nw, "C chunk col"));
            BRMC(chunks.alloc(&cv, nw, "C chunk val"));
            ch_col.push_back(cc);
            ch_val.push_back(cv);
            ch_n.push_back(nw);
        }
        clk.begin(P_COPY);
        HIPC(launch_compact(crpt_w, aw.M, aw, o.ccol, o.cval, cc, cv, nw, s));
        dev_flush_deferred();
    }
    // C-bar is no longer needed: return it before C is allocated.
    HIPC(hipStreamSynchronize(s));
    sc.free_now(bcol_bar);
    sc.free_now(bval_bar);
    clk.begin(P_PNNZ);
    HIPC(scan_i64(p.rowcnt, M, p.crpt, s));
    int64_t nnzC = 0;
    BRMC(d2h(&nnzC, p.crpt + M, sizeof(int64_t), s));
    if (nnzC > max_nnz) {
        set_err("nnz(C) = %lld exceeds the int32 row pointer", (long long)nnzC);
        return BRM_E_OVERFLOW;
    }
    if (vcol) {
        if (voff != nnzC) {
            set_err("Upper row waves: %lld entries compacted, nnz(C) = %lld", (long long)voff, (long long)nnzC);
            return BRM_E_INTERNAL;
        }
        HIPC(hipStreamSynchronize(s));
        *ccol_out = vcol;
        *cval_out = vval;
        *nnz_out = nnzC;
        return BRM_OK;
    }
    clk.begin(P_ALLOC);
    BRMC(sc.alloc(ccol_out, nnzC, "C col"));
    BRMC(sc.alloc(cval_out, nnzC, "C val"));
    clk.begin(P_COPY);
    int64_t off = 0;
    for (size_t w = 0; w < ch_n.size(); ++w) {
        if (ch_n[w] > 0) {
            HIPC(hipMemcpyAsync(*ccol_out + off, ch_col[w], sizeof(int) * ch_n[w], hipMemcpyDeviceToDevice, s));
            HIPC(hipMemcpyAsync(*cval_out + off, ch_val[w], sizeof(double) * ch_n[w], hipMemcpyDeviceToDevice, s));
        }
        off += ch_n[w];
    }
    if (off != nnzC) {
        set_err("Upper row waves: chunk sizes (%lld) disagree with nnz(C) (%lld)", (long long)off, (long long)nnzC);
        return BRM_E_INTERNAL;
    }
    HIPC(hipStreamSynchronize(s));
    *nnz_out = nnzC;
    return BRM_OK;
}
}  // namespace

namespace {
// Products of at most SMALL_MAX_ROWS rows whose rows all have at most `group`
// (32 or 64) lists and products: one fill, one kernel and one host round trip
// (small.hip).  C is allocated for `group` entries per row (a bound of nnz(C)
// when every row is short), so nothing waits on the device before the end.
// *handled = false when a row is longer (nothing is kept).
int run_small_g(const DevIn& in, int group, int64_t max_nnz, hipStream_t s, PhaseClock& clk, DevOut* out,
                bool* handled) {
    *handled = false;
    const int M = in.M;
    const int64_t cap = (int64_t)M * group;
    if (cap > max_nnz) return BRM_OK;
    Scratch sc;
    SmallState* st;
    int64_t* crpt;
    int* ccol;
    double* cval;
    clk.begin(P_PRE);
    BRMC(sc.alloc(reinterpret_cast<int64_t**>(&st), small_state_words(M, group), "small-product state"));
    BRMC(sc.alloc(&crpt, (int64_t)M + 1, "crpt"));
    BRMC(sc.alloc(&ccol, cap, "C col"));
    BRMC(sc.alloc(&cval, cap, "C val"));
    RowArgs a{};
    a.arpt = in.arpt;
    a.acol = in.acol;
    a.aval = in.aval;
    a.brpt = in.brpt;
    a.bcol = in.bcol;
    a.bval = in.bval;
    a.M = M;
    a.K = in.K;
    a.N = in.N;
    clk.begin(P_COMPUTE);
    HIPC(launch_small(a, in.nnzA, in.nnzB, crpt, ccol, cval, st, group, s));
    SmallState h{};
    BRMC(d2h(&h, st, sizeof(h), s));
    if (h.err & DERR_INVALID) {
        set_err("validation: malformed CSR input (row pointers, column range or unsorted/duplicate B columns)");
        return BRM_E_INVALID;
    }
    if (h.err) return dev_err_status(h.err, "small product");
    if (h.miss) return BRM_OK;
    int64_t nnz = 0, ipt = 0;
    for (int i = 0; i < SMALL_SLOTS; ++i) {
        nnz += h.nnz[i];
        ipt += h.ip[i];
    }
    sc.release(crpt);
    sc.release(ccol);
    sc.release(cval);
    out->crpt = crpt;
    out->ccol = ccol;
    out->cval = cval;
    out->nnz = nnz;
    brm_stats& ts = thread_stats();
    ts.total_ip = ipt;
    ts.nnz_c = nnz;
    ts.rows_by_tier[BIN_SHORT] = M;
    *handled = true;
    return BRM_OK;
}

// Groups of 32 lanes (two rows per wavefront) when the average row has at
// most 32 products, else 64; a row too long for 32 lanes retries at 64.
int run_small(const DevIn& in, double avg_ip, int64_t max_nnz, hipStream_t s, PhaseClock& clk, DevOut* out,
              bool* handled) {
    *handled = false;
    if (in.nnzA > INT32_MAX || in.nnzB > INT32_MAX) return BRM_OK;
    if (in.M < 1 || in.K < 1 || in.nnzA < 1 || in.nnzB < 1) return BRM_OK;  // C is empty: the general pipeline
    if (avg_ip <= 32.0) {
        BRMC(run_small_g(in, 32, max_nnz, s, clk, out, handled));
        if (*handled) return BRM_OK;
    }
    return run_small_g(in, 64, max_nnz, s, clk, out, handled);
}
}  // namespace

int run_spgemm(const DevIn& in, bool upper, int64_t max_nnz, hipStream_t s, DevOut* out, double* ph,
               const ExtOut* ext) {
    PhaseClock clk(s, ph != nullptr);
    memset(&t_stats, 0, sizeof(t_stats));
    HostTrace tr(s);
    tr.mark("spgemm start", in.M);
    Pattern p;
    if (ext && upper) {
        set_err("caller-owned outputs are for BRMerge-Precise");
        return BRM_E_INVALID;
    }
    if (ext) p.crpt_ext = ext->crpt;
    // the fused path when every row is likely short: lists per row times
    // entries per B row at most 64 on average.  It serves BRMerge-Upper as
    // well: it has no symbolic phase, and it writes C into a buffer sized by
    // an upper bound (G entries per row), compacted as it goes.
    const double avg_l = in.M > 0 ? (double)in.nnzA / in.M : 0.0, avg_b = in.K > 0 ? (double)in.nnzB / in.K : 0.0;
    if (!ext && params().small_fused && in.M > 0 && in.M <= SMALL_MAX_ROWS && !params().force_long &&
        params().hash_cap == 0 && avg_l * avg_b <= 64.0) {
        bool handled = false;
        BRMC(run_small(in, avg_l * avg_b, max_nnz, s, clk, out, &handled));
        if (handled) {
            clk.begin(P_CLEAN);
            HIPC(hipStreamSynchronize(s));
            clk.end();
            if (ph) clk.collect(ph);
            dev_flush_deferred();
            return BRM_OK;
        }
    }
    BRMC(pattern_prepare(in, s, &clk, p));
    const int M = in.M;
    for (int b = 0; b < 4; ++b) t_stats.rows_by_tier[b] = p.bc[b];
    t_stats.rows_by_tier[BIN_MEDIUM] += p.bc[BIN_WMED] + p.bc[BIN_WMED2] + p.bc[BIN_WMED3] + p.bc[BIN_WMED4];
    t_stats.rows_by_tier[BIN_SHORT] += p.bc[BIN_TINY];
    t_stats.rows_by_tier[BIN_MEDIUM] += p.bc[BIN_HASH] + p.bc[BIN_HASH2];
    t_stats.total_ip = p.total_ip;
    Scratch& sc = p.sc;
    int* ccol = nullptr;
    double* cval = nullptr;
    int64_t nnzC = 0;

    if (!upper) {
        // Small products size C by the product count (an upper bound of
        // nnz(C), at most SMALL_IP elements) and read nnz(C) with the
        // numeric phase's error flag: one host round trip less.
        // Caller-owned C: the short-row stash is not used (it would take
        // scratch memory), and the product count bounds nnz(C) only when the
        // caller's capacity covers it.
        constexpr int64_t SMALL_IP = (int64_t)1 << 24;
        const bool small = p.total_ip <= SMALL_IP && p.total_ip <= max_nnz && (!ext || p.total_ip <= ext->cap);
        BRMC(pattern_symbolic(p, s, &clk, ext == nullptr, !small));
        nnzC = p.nnzC;
        if (!small && nnzC > max_nnz) {
            set_err("nnz(C) = %lld exceeds the int32 row pointer", (long long)nnzC);
            return BRM_E_OVERFLOW;
        }
        const int64_t cap = small ? p.total_ip : nnzC;
        clk.begin(P_ALLOC);
        if (ext) {
            if (!small && nnzC > ext->cap) {
                if (ext->nnz_needed) *ext->nnz_needed = nnzC;
                set_err("nnz(C) = %lld exceeds the caller's capacity of %lld entries", (long long)nnzC,
                        (long long)ext->cap);
                return BRM_E_OVERFLOW;
            }
            ccol = ext->ccol;
            cval = ext->cval;
        } else {
            BRMC(sc.alloc(&ccol, cap, "C col"));
            BRMC(sc.alloc(&cval, cap, "C val"));
        }
        tr.mark("C allocated", cap);
        clk.begin(P_COMPUTE);
        // With long rows, the short and medium tiers run on a side stream,
        // beside the long-row engine: they fill the CUs its passes leave idle
        // (planning round trips, the last passes of the heaviest rows).
        // Different rows, disjoint parts of C.
        SideStream side(s, p.bc[BIN_LONG] > 0 && !ext);
        BRMC(side.fork());
        BRMC(pattern_numeric_tiers(p, ccol, cval, side.stream()));
        BRMC(side.mark());
        tr.mark("short+medium numeric");
        BRMC(pattern_numeric_long(p, ccol, cval, s));
        BRMC(side.join());
        tr.mark("long numeric");
        if (small) {
            int e = 0;
            BRMC(d2h_parts({{&e, p.derr, sizeof(int)}, {&nnzC, p.crpt + M, sizeof(int64_t)}}, s));
            BRMC(dev_err_status(e, "numeric"));
        } else {
            BRMC(check_dev_err(p.derr, s, "numeric"));
        }
    } else {
        BRMC(run_upper(p, max_nnz, s, clk, &ccol, &cval, &nnzC));
    }
    clk.begin(P_CLEAN);
    HIPC(hipStreamSynchronize(s));
    clk.end();
    HIPC(hipStreamSynchronize(s));
    if (ph) clk.collect(ph);

    t_stats.nnz_c = nnzC;
    sc.release(p.crpt);
    sc.release(ccol);
    sc.release(cval);
    out->crpt = p.crpt;
    out->ccol = ccol;
    out->cval = cval;
    out->nnz = nnzC;
    dev_flush_deferred();
    return BRM_OK;
}

namespace {

void fill_precise(brm_precise_timing* t, const double* ph, double total, double h2d, double d2h_) {
    if (!t) return;
    t->pre_allocate = ph[P_PRE];
    t->compute_flop = ph[P_FLOP];
    t->prefix_sum_flop = ph[P_PFLOP];
    t->load_balance = ph[P_LB];
    t->symbolic = ph[P_SYM];
    t->prefix_sum_nnz = ph[P_PNNZ];
    t->allocate_C = ph[P_ALLOC];
    t->compute = ph[P_COMPUTE];
    t->cleanup = ph[P_CLEAN];
    t->total = total;
    t->h2d = h2d;
    t->d2h = d2h_;
}

void fill_upper(brm_upper_timing* t, const double* ph, double total, double h2d, double d2h_) {
    if (!t) return;
    t->thread = 0;
    t->pre_allocate = ph[P_PRE];
    t->compute_flop = ph[P_FLOP];
    t->prefix_sum_flop = ph[P_PFLOP];
    t->load_balance = ph[P_LB];
    t->allocate_hook = 0;
    t->compute = ph[P_COMPUTE];
    t->prefix_sum_nnz = ph[P_PNNZ];
    t->allocate_C = ph[P_ALLOC];
    t->copy = ph[P_COPY];
    t->cleanup = ph[P_CLEAN];
    t->total = total;
    t->h2d = h2d;
    t->d2h = d2h_;
}

// Host-pointer entry shared by the four host variants.
template <typename RptT>
int host_entry(const int* arpt, const int* acol, const double* aval, const int* brpt, const int* bcol,
               const double* bval, RptT** crpt_, int** ccol_, double** cval_, int M, int K, int N, RptT* cnnz_,
               bool upper, double* ph, double* total, double* h2d, double* d2h_) {
    if (!crpt_ || !ccol_ || !cval_ || !cnnz_) {
        set_err("null output pointer");
        return BRM_E_INVALID;
    }
    if (M < 0 || K < 0 || N < 0 || (M > 0 && !arpt) || (K > 0 && !brpt)) {
        set_err("invalid dimensions or null row pointer");
        return BRM_E_INVALID;
    }
    const double t0 = now_s();
    BRMC(use_device());
    const int64_t nnzA = M > 0 ? arpt[M] : 0;
    const int64_t nnzB = K > 0 ? brpt[K] : 0;
    if (nnzA < 0 || nnzB < 0 || (nnzA > 0 && (!acol || !aval)) || (nnzB > 0 && (!bcol || !bval))) {
        set_err("invalid nnz or null column/value pointer");
        return BRM_E_INVALID;
    }
    hipStream_t s = nullptr;
    Scratch sc;
    int *darpt, *dacol, *dbrpt, *dbcol;
    double *daval, *dbval;
    BRMC(sc.alloc(&darpt, (int64_t)M + 1, "A rpt"));
    BRMC(sc.alloc(&dacol, nnzA, "A col"));
    BRMC(sc.alloc(&daval, nnzA, "A val"));
    BRMC(sc.alloc(&dbrpt, (int64_t)K + 1, "B rpt"));
    BRMC(sc.alloc(&dbcol, nnzB, "B col"));
    BRMC(sc.alloc(&dbval, nnzB, "B val"));
    if (M == 0) {
        int z = 0;
        HIPC(hipMemcpyAsync(darpt, &z, sizeof(int), hipMemcpyHostToDevice, s));
    } else {
        HIPC(hipMemcpyAsync(darpt, arpt, sizeof(int) * ((size_t)M + 1), hipMemcpyHostToDevice, s));
    }
    if (K == 0) {
        int z = 0;
        HIPC(hipMemcpyAsync(dbrpt, &z, sizeof(int), hipMemcpyHostToDevice, s));
    } else {
        HIPC(hipMemcpyAsync(dbrpt, brpt, sizeof(int) * ((size_t)K + 1), hipMemcpyHostToDevice, s));
    }
    if (nnzA > 0) {
        HIPC(hipMemcpyAsync(dacol, acol, sizeof(int) * nnzA, hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(daval, aval, sizeof(double) * nnzA, hipMemcpyHostToDevice, s));
    }
    if (nnzB > 0) {
        HIPC(hipMemcpyAsync(dbcol, bcol, sizeof(int) * nnzB, hipMemcpyHostToDevice, s));
        HIPC(hipMemcpyAsync(dbval, bval, sizeof(double) * nnzB, hipMemcpyHostToDevice, s));
    }
    HIPC(hipStreamSynchronize(s));
    const double t1 = now_s();

    DevIn in{darpt, dacol, daval, nnzA, dbrpt, dbcol, dbval, nnzB, M, K, N};
    DevOut out{};
    const int64_t max_nnz = sizeof(RptT) == 4 ? (int64_t)INT32_MAX : INT64_MAX;
    BRMC(run_spgemm(in, upper, max_nnz, s, &out, ph));
    Scratch outs;
    outs.bufs = {out.crpt, out.ccol, out.cval};

    const double t2 = now_s();
    RptT* hrpt = new (std::nothrow) RptT[(size_t)M + 1];
    int* hcol = new (std::nothrow) int[(size_t)(out.nnz > 0 ? out.nnz : 1)];
    double* hval = new (std::nothrow) double[(size_t)(out.nnz > 0 ? out.nnz : 1)];
    if (!hrpt || !hcol || !hval) {
        delete[] hrpt;
        delete[] hcol;
        delete[] hval;
        set_err("host allocation of C failed");
        return BRM_E_NOMEM;
    }
    int st = BRM_OK;
    if (sizeof(RptT) == 8) {
        st = d2h(hrpt, out.crpt, sizeof(int64_t) * ((size_t)M + 1), s);
    } else {
        int* tmp;
        st = sc.alloc(&tmp, (int64_t)M + 1, "crpt32");
        if (st == BRM_OK) st = hip_status(launch_rpt_to_i32(out.crpt, (int64_t)M + 1, tmp, s), "crpt32");
        if (st == BRM_OK) st = d2h(hrpt, tmp, sizeof(int) * ((size_t)M + 1), s);
    }
    if (st == BRM_OK && out.nnz > 0) {
        st = d2h(hcol, out.ccol, sizeof(int) * out.nnz, s);
        if (st == BRM_OK) st = d2h(hval, out.cval, sizeof(double) * out.nnz, s);
    }
    if (st != BRM_OK) {
        delete[] hrpt;
        delete[] hcol;
        delete[] hval;
        return st;
    }
    const double t3 = now_s();
    *crpt_ = hrpt;
    *ccol_ = hcol;
    *cval_ = hval;
    *cnnz_ = (RptT)out.nnz;
    *h2d = t1 - t0;
    *d2h_ = t3 - t2;
    *total = t3 - t0;
    dev_flush_deferred();
    return BRM_OK;
}

template <typename Tm>
int dev_entry(const int* arpt, const int* acol, const double* aval, int64_t nnzA, const int* brpt, const int* bcol,
              const double* bval, int64_t nnzB, int M, int K, int N, int64_t** crpt, int** ccol, double** cval,
              int64_t* cnnz, bool upper, Tm* timing, void* stream) {
    if (!crpt || !ccol || !cval || !cnnz || (M > 0 && !arpt) || (K > 0 && !brpt)) {
        set_err("null pointer argument");
        return BRM_E_INVALID;
    }
    const double t0 = now_s();
    BRMC(use_device());
    DevIn in{arpt, acol, aval, nnzA, brpt, bcol, bval, nnzB, M, K, N};
    DevOut out{};
    double ph[P_N] = {0};
    BRMC(run_spgemm(in, upper, INT64_MAX, (hipStream_t)stream, &out, ph));
    *crpt = out.crpt;
    *ccol = out.ccol;
    *cval = out.cval;
    *cnnz = out.nnz;
    const double total = now_s() - t0;
    if constexpr (std::is_same<Tm, brm_precise_timing>::value) fill_precise(timing, ph, total, 0, 0);
    else fill_upper(timing, ph, total, 0, 0);
    return BRM_OK;
}

}  // namespace
}  // namespace brm

using namespace brm;

extern "C" {

int brm_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        set_err("no HIP device %d", device);
        return BRM_E_DEVICE;
    }
    t_device = device;
    BRMC(hip_status(hipSetDevice(device), "hipSetDevice"));
    return noalloc_setup(device);  // the fixed scratch of brm_spgemm_precise_dev_noalloc on this device
}

const char* brm_last_error(void) { return t_err.c_str(); }

const char* brm_version(void) { return "brmerge-mi355x 0.1 (gfx950)"; }

int brm_spgemm_precise_i32(const int* arpt, const int* acol, const double* aval, const int* brpt, const int* bcol,
                           const double* bval, int** crpt, int** ccol, double** cval, int M, int K, int N, int* cnnz,
                           brm_precise_timing* timing) {
    double ph[P_N] = {0}, tot = 0, h2d = 0, d2h_ = 0;
    int st = host_entry<int>(arpt, acol, aval, brpt, bcol, bval, crpt, ccol, cval, M, K, N, cnnz, false, ph, &tot,
                             &h2d, &d2h_);
    if (st == BRM_OK) fill_precise(timing, ph, tot, h2d, d2h_);
    dev_flush_deferred();
    return st;
}

int brm_spgemm_upper_i32(const int* arpt, const int* acol, const double* aval, const int* brpt, const int* bcol,
                         const double* bval, int** crpt, int** ccol, double** cval, int M, int K, int N, int* cnnz,
                         brm_upper_timing* timing) {
    double ph[P_N] = {0}, tot = 0, h2d = 0, d2h_ = 0;
    int st = host_entry<int>(arpt, acol, aval, brpt, bcol, bval, crpt, ccol, cval, M, K, N, cnnz, true, ph, &tot,
                             &h2d, &d2h_);
    if (st == BRM_OK) fill_upper(timing, ph, tot, h2d, d2h_);
    dev_flush_deferred();
    return st;
}

int brm_spgemm_precise_i64(const int* arpt, const int* acol, const double* aval, const int* brpt, const int* bcol,
                           const double* bval, int64_t** crpt, int** ccol, double** cval, int M, int K, int N,
                           int64_t* cnnz, brm_precise_timing* timing) {
    double ph[P_N] = {0}, tot = 0, h2d = 0, d2h_ = 0;
    int st = host_entry<int64_t>(arpt, acol, aval, brpt, bcol, bval, crpt, ccol, cval, M, K, N, cnnz, false, ph,
                                 &tot, &h2d, &d2h_);
    if (st == BRM_OK) fill_precise(timing, ph, tot, h2d, d2h_);
    dev_flush_deferred();
    return st;
}

int brm_spgemm_upper_i64(const int* arpt, const int* acol, const double* aval, const int* brpt, const int* bcol,
                         const double* bval, int64_t** crpt, int** ccol, double** cval, int M, int K, int N,
                         int64_t* cnnz, brm_upper_timing* timing) {
    double ph[P_N] = {0}, tot = 0, h2d = 0, d2h_ = 0;
    int st = host_entry<int64_t>(arpt, acol, aval, brpt, bcol, bval, crpt, ccol, cval, M, K, N, cnnz, true, ph,
                                 &tot, &h2d, &d2h_);
    if (st == BRM_OK) fill_upper(timing, ph, tot, h2d, d2h_);
    dev_flush_deferred();
    return st;
}

void brm_host_free_i32(int* p) { delete[] p; }
void brm_host_free_i64(int64_t* p) { delete[] p; }
void brm_host_free_f64(double* p) { delete[] p; }

int brm_spgemm_precise_dev(const int* arpt, const int* acol, const double* aval, int64_t nnzA, const int* brpt,
                           const int* bcol, const double* bval, int64_t nnzB, int M, int K, int N, int64_t** crpt,
                           int** ccol, double** cval, int64_t* cnnz, brm_precise_timing* timing, void* stream) {
    int st = dev_entry(arpt, acol, aval, nnzA, brpt, bcol, bval, nnzB, M, K, N, crpt, ccol, cval, cnnz, false,
                       timing, stream);
    dev_flush_deferred();
    return st;
}

int brm_spgemm_upper_dev(const int* arpt, const int* acol, const double* aval, int64_t nnzA, const int* brpt,
                         const int* bcol, const double* bval, int64_t nnzB, int M, int K, int N, int64_t** crpt,
                         int** ccol, double** cval, int64_t* cnnz, brm_upper_timing* timing, void* stream) {
    int st = dev_entry(arpt, acol, aval, nnzA, brpt, bcol, bval, nnzB, M, K, N, crpt, ccol, cval, cnnz, true,
                       timing, stream);
    dev_flush_deferred();
    return st;
}

void* brm_dev_malloc(uint64_t bytes) {
    if (use_device() != BRM_OK) return nullptr;
    return dev_alloc((size_t)bytes);
}
void brm_dev_free(void* p) { dev_free(p); }

int brm_rpt_to_i32(const int64_t* in, int64_t n, int* out, void* stream) {
    if (n < 0 || (n > 0 && (!in || !out))) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    if (n == 0) return BRM_OK;
    BRMC(use_device());
    hipStream_t s = (hipStream_t)stream;
    int64_t last = 0;
    BRMC(d2h(&last, in + n - 1, sizeof(int64_t), s));
    if (last > INT32_MAX) {
        set_err("row pointer value %lld exceeds int32", (long long)last);
        return BRM_E_OVERFLOW;
    }
    HIPC(launch_rpt_to_i32(in, n, out, s));
    HIPC(hipStreamSynchronize(s));
    return BRM_OK;
}
void brm_dev_trim(void) { dev_trim(); }

int brm_flop_count_dev(const int* arpt, const int* acol, int64_t nnzA, const int* brpt, int M, int K,
                       int64_t* row_ip, int64_t* total, void* stream) {
    if (!total || (M > 0 && !arpt) || (K > 0 && !brpt) || M < 0 || nnzA < 0) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    hipStream_t s = (hipStream_t)stream;
    Scratch sc;
    int64_t* ipA;
    BRMC(sc.alloc(&ipA, nnzA + 1, "ipA"));
    HIPC(scan_alen(acol, brpt, K, nnzA, ipA, s));
    if (row_ip && M > 0) {
        RowArgs a{};
        a.arpt = arpt;
        a.ipA = ipA;
        a.M = M;
        // row_ip[i] = ipA[arpt[i+1]] - ipA[arpt[i]] via the row-stat kernel's
        // helper: reuse the long-info kernel over all rows.
        int* rows;
        int* lens;
        BRMC(sc.alloc(&rows, M, "rows"));
        BRMC(sc.alloc(&lens, M, "lens"));
        std::vector<int> ids(M);
        for (int i = 0; i < M; ++i) ids[i] = i;
        HIPC(hipMemcpyAsync(rows, ids.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
        HIPC(launch_long_info(rows, M, a, row_ip, lens, s));
    }
    int st = d2h(total, ipA + nnzA, sizeof(int64_t), s);
    dev_flush_deferred();
    return st;
}

int brm_flop_count(const int* arpt, const int* acol, const int* brpt, int M, int K, int64_t* total) {
    if (!total || M < 0 || K < 0 || (M > 0 && !arpt) || (K > 0 && !brpt)) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    const int64_t nnzA = M > 0 ? arpt[M] : 0;
    Scratch sc;
    int *dacol, *dbrpt;
    BRMC(sc.alloc(&dacol, nnzA, "A col"));
    BRMC(sc.alloc(&dbrpt, (int64_t)K + 1, "B rpt"));
    hipStream_t s = nullptr;
    if (nnzA > 0) HIPC(hipMemcpyAsync(dacol, acol, sizeof(int) * nnzA, hipMemcpyHostToDevice, s));
    if (K > 0) HIPC(hipMemcpyAsync(dbrpt, brpt, sizeof(int) * ((size_t)K + 1), hipMemcpyHostToDevice, s));
    for (int64_t j = 0; j < nnzA; ++j) {
        if (acol[j] < 0 || acol[j] >= K) {
            set_err("A column index out of range");
            return BRM_E_INVALID;
        }
    }
    return brm_flop_count_dev(nullptr, dacol, nnzA, dbrpt, 0, K, nullptr, total, s);
}

int brm_set_param(const char* name, int64_t value) {
    if (!name) return BRM_E_INVALID;
    Params& p = params();
    if (!strcmp(name, "ws_budget_bytes")) { if (value < 0) return BRM_E_INVALID; p.ws_budget_bytes = value; }
    else if (!strcmp(name, "med_cap")) { if (value < SHORT_MAX || value > MED_CAP) return BRM_E_INVALID; p.med_cap = (int)value; }
    else if (!strcmp(name, "leaf_cap")) { if (value < 1 || value > MED_CAP) return BRM_E_INVALID; p.leaf_cap = (int)value; }
    else if (!strcmp(name, "force_long")) { p.force_long = value ? 1 : 0; }
    else if (!strcmp(name, "profile")) { p.profile = value ? 1 : 0; }
    else if (!strcmp(name, "lds_cfg")) { if (value < 0 || value >= N_LDS_CFG) return BRM_E_INVALID; p.lds_cfg = (int)value; }
    else if (!strcmp(name, "group_cfg")) { if (value < 0 || value >= N_GROUP_CFG) return BRM_E_INVALID; p.group_cfg = (int)value; }
    else if (!strcmp(name, "wmed_cap")) { if (value < -1 || value > WMED_CAP_BITMAP) return BRM_E_INVALID; p.wmed_cap = (int)value; }
    else if (!strcmp(name, "launch_chunk")) { if (value < 0) return BRM_E_INVALID; p.launch_chunk = value; }
    else if (!strcmp(name, "wmed_engine")) { if (value < 0 || value > 1) return BRM_E_INVALID; p.wmed_engine = (int)value; }
    else if (!strcmp(name, "cbar_budget_bytes")) { if (value < 0) return BRM_E_INVALID; p.cbar_budget_bytes = value; }
    else if (!strcmp(name, "upper_vm")) { p.upper_vm = value ? 1 : 0; }
    else if (!strcmp(name, "small_fused")) { p.small_fused = value ? 1 : 0; }
    else if (!strcmp(name, "wmed_count")) { p.wmed_count = value ? 1 : 0; }
    else if (!strcmp(name, "wmed_bitmap")) { p.wmed_bitmap = value ? 1 : 0; }
    else if (!strcmp(name, "hash_cap")) { if (value < 0 || value > HASH_CAP) return BRM_E_INVALID; p.hash_cap = (int)value; }
    else if (!strcmp(name, "sym_chunk_products")) { if (value < 1 || value > ((int64_t)1 << 30)) return BRM_E_INVALID; p.sym_chunk_products = value; }
    else return BRM_E_INVALID;
    return BRM_OK;
}

int64_t brm_get_param(const char* name) {
    if (!name) return -1;
    Params& p = params();
    if (!strcmp(name, "ws_budget_bytes")) return p.ws_budget_bytes;
    if (!strcmp(name, "med_cap")) return p.med_cap;
    if (!strcmp(name, "leaf_cap")) return p.leaf_cap;
    if (!strcmp(name, "force_long")) return p.force_long;
    if (!strcmp(name, "profile")) return p.profile;
    if (!strcmp(name, "lds_cfg")) return p.lds_cfg;
    if (!strcmp(name, "group_cfg")) return p.group_cfg;
    if (!strcmp(name, "wmed_cap")) return p.wmed_cap;
    if (!strcmp(name, "launch_chunk")) return p.launch_chunk;
    if (!strcmp(name, "wmed_engine")) return p.wmed_engine;
    if (!strcmp(name, "cbar_budget_bytes")) return p.cbar_budget_bytes;
    if (!strcmp(name, "sym_chunk_products")) return p.sym_chunk_products;
    if (!strcmp(name, "hash_cap")) return p.hash_cap;
    if (!strcmp(name, "upper_vm")) return p.upper_vm;
    if (!strcmp(name, "small_fused")) return p.small_fused;
    if (!strcmp(name, "wmed_count")) return p.wmed_count;
    if (!strcmp(name, "wmed_bitmap")) return p.wmed_bitmap;
    return -1;
}

void brm_debug_fail_alloc_after(int64_t n) { dev_set_fail_after(n); }

int brm_bench_copy(void* dst, const void* src, uint64_t bytes, void* stream) {
    if (!dst || !src || (bytes & 15)) return BRM_E_INVALID;
    BRMC(use_device());
    HIPC(launch_copy(dst, src, bytes, (hipStream_t)stream));
    HIPC(hipStreamSynchronize((hipStream_t)stream));
    return BRM_OK;
}

int brm_checksum_dev(const void* p, uint64_t bytes, uint64_t* out, void* stream) {
    if (!out || (bytes & 3) || (bytes > 0 && !p)) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    hipStream_t s = (hipStream_t)stream;
    Scratch sc;
    uint64_t* d;
    BRMC(sc.alloc(&d, 1, "checksum"));
    HIPC(launch_checksum(p, (int64_t)(bytes / 4), d, s));
    return d2h(out, d, sizeof(uint64_t), s);
}

int brm_last_stats(brm_stats* out) {
    if (!out) return BRM_E_INVALID;
    *out = t_stats;
    return BRM_OK;
}

int brm_memcpy_h2d(void* dst_dev, const void* src_host, uint64_t bytes) {
    if (bytes == 0) return BRM_OK;
    if (!dst_dev || !src_host) return BRM_E_INVALID;
    BRMC(use_device());
    HIPC(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice));
    return BRM_OK;
}

int brm_memcpy_d2h(void* dst_host, const void* src_dev, uint64_t bytes) {
    if (bytes == 0) return BRM_OK;
    if (!dst_host || !src_dev) return BRM_E_INVALID;
    BRMC(use_device());
    HIPC(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost));
    return BRM_OK;
}

}  // extern "C"
