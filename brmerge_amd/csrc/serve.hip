// Request server: B operands resident in HBM, A * B requests admitted by an
// HBM budget and batched as they arrive.
//
// Every registered B is appended to one stacked B (rows of B_0, then B_1,
// ...), so a batch of requests against any mix of B operands is a single
// fused product: the requests' A matrices stacked by rows, each A's columns
// shifted to its B's rows in the stack.  Every request's result is the
// product it would get alone, bit for bit (each C row depends only on its
// own A row and the B rows it selects).
//
// Admission: a request reserves its BRMerge-Upper bound, 12 bytes per
// intermediate product plus its row pointer, 12*ip + 8*(M+1), against the
// server budget; the reservation shrinks to the result's size when the
// batch finishes and is returned when the caller releases the result.
// Each worker's long-row workspace is bounded too: an eighth of the budget,
// split evenly among the workers, is their per-call workspace budget (on top
// of the admitted results), so concurrent batches never size workspace by
// reading the device's free memory at the same time.
// Worker threads (one HIP stream each) take every admissible queued request
// (FIFO, up to max_batch) whenever they are free, so batches formed while
// others run overlap on the device.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "brm_device.h"
#include "pipeline.h"

namespace brm {
namespace {

struct ResidentB {
    int64_t k0;  // first row in the stack
    int64_t nnz;
    int K, N;
};

// Device arrays shared by the results of one batch.
struct BatchMem {
    std::vector<void*> bufs;
    ~BatchMem() {
        for (void* p : bufs) dev_free(p);
    }
};

struct Request {
    int64_t id = 0;
    int b = 0;
    int N = 0;  // columns of B (copied at submit: the B table may grow meanwhile)
    const int* arpt = nullptr;
    const int* acol = nullptr;
    const double* aval = nullptr;
    int64_t nnzA = 0;
    int M = 0;
    int64_t ip = 0;
    int64_t reserved = 0;  // bytes held against the budget
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    int status = BRM_OK;
    std::string err;
    int batch_size = 0;
    std::shared_ptr<BatchMem> mem;
    const int64_t* crpt = nullptr;
    const int* ccol = nullptr;
    const double* cval = nullptr;
    int64_t nnz = 0;
};

}  // namespace
}  // namespace brm

using namespace brm;

struct brm_server {
    int device = 0;
    int64_t budget = 0;     // admission budget (results)
    int64_t ws_share = 0;   // long-row workspace budget of each worker
    int max_batch = 64;
    std::vector<ResidentB> bs;
    // the stacked B (device), owned
    std::vector<void*> stack_bufs;
    DevIn stack{};

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
    int64_t next_id = 1;
    int64_t in_use = 0;
    int inflight = 0;
    bool rebuilding = false;
    bool stop = false;
    std::vector<std::thread> workers;

    std::mutex aux_mu;  // submit-side validation and product counts
    hipStream_t aux = nullptr;

    // statistics of completed requests
    std::vector<double> latency;
    int64_t batches = 0, batched = 0, total_ip = 0;
    double first_submit = -1, last_done = 0;

    ~brm_server() {
        for (void* p : stack_bufs) dev_free(p);
        if (aux) (void)hipStreamDestroy(aux);
    }
};

namespace {

void finish_batch(brm_server* S, std::vector<std::shared_ptr<Request>>& batch, int st, const std::string& err) {
    std::lock_guard<std::mutex> lk(S->mu);
    const double t = now_s();
    for (auto& r : batch) {
        r->done = true;
        r->t_done = t;
        r->status = st;
        r->err = err;
        r->batch_size = (int)batch.size();
        const int64_t keep = st == BRM_OK ? 12 * r->nnz + 8 * ((int64_t)r->M + 1) : 0;
        S->in_use -= r->reserved - std::min(keep, r->reserved);
        r->reserved = std::min(keep, r->reserved);
        if (st == BRM_OK) {
            S->latency.push_back(r->t_done - r->t_submit);
            S->total_ip += r->ip;
            S->last_done = t;
        }
    }
    S->batches += 1;
    S->batched += (int64_t)batch.size();
    S->inflight -= 1;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
}

// Runs one batch on stream s; fills the requests' result views.
int run_batch(brm_server* S, std::vector<std::shared_ptr<Request>>& batch, hipStream_t s) {
    const int n = (int)batch.size();
    std::vector<brm_pair> pairs(n);
    std::vector<int64_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const Request& r = *batch[i];
        const ResidentB& B = S->bs[r.b];
        pairs[i] = brm_pair{r.arpt, r.acol, r.aval, r.nnzA, S->stack.brpt + B.k0, S->stack.bcol, S->stack.bval,
                            B.nnz, r.M, B.K, B.N};
        off[i] = B.k0;
    }
    FusedOut fo;
    BRMC(fuse_pairs(pairs.data(), n, off.data(), s, &fo));
    DevIn in = fo.in;
    in.brpt = S->stack.brpt;
    in.bcol = S->stack.bcol;
    in.bval = S->stack.bval;
    in.nnzB = S->stack.nnzB;
    in.K = S->stack.K;
    in.N = S->stack.N;
    in.b_checked = 1;
    DevOut out{};
    BRMC(run_spgemm(in, false, INT64_MAX, s, &out, nullptr));
    auto mem = std::make_shared<BatchMem>();
    mem->bufs = {out.ccol, out.cval};
    Scratch sc;
    sc.bufs.push_back(out.crpt);
    // per-request row pointers
    std::vector<int64_t> ebase(n + 1, 0);
    for (int i = 0; i < n; ++i) ebase[i + 1] = ebase[i] + batch[i]->M + 1;
    int64_t* rpt_all;
    int64_t *tab_d, *base_d;
    rpt_all = (int64_t*)dev_alloc(sizeof(int64_t) * (size_t)std::max<int64_t>(ebase[n], 1));
    if (!rpt_all) {
        set_err("device allocation of request row pointers failed");
        return BRM_E_NOMEM;
    }
    mem->bufs.push_back(rpt_all);
    BRMC(sc.alloc(&tab_d, 2 * ((int64_t)n + 1), "request tables"));
    BRMC(sc.alloc(&base_d, n, "request bases"));
    std::vector<int64_t> tab(fo.mbase);
    tab.insert(tab.end(), ebase.begin(), ebase.end());
    HIPC(hipMemcpyAsync(tab_d, tab.data(), sizeof(int64_t) * tab.size(), hipMemcpyHostToDevice, s));
    HIPC(launch_split_rpt(out.crpt, tab_d, tab_d + (n + 1), n, ebase[n], rpt_all, base_d, s));
    std::vector<int64_t> base(n + 1);
    BRMC(d2h(base.data(), base_d, sizeof(int64_t) * n, s));
    base[n] = out.nnz;
    for (int i = 0; i < n; ++i) {
        Request& r = *batch[i];
        r.mem = mem;
        r.crpt = rpt_all + ebase[i];
        r.ccol = out.ccol + base[i];
        r.cval = out.cval + base[i];
        r.nnz = base[i + 1] - base[i];
    }
    return BRM_OK;
}

void worker_main(brm_server* S) {
    set_thread_device(S->device);
    set_thread_ws_budget(S->ws_share);
    set_thread_side_streams(false);  // the workers' own streams run beside each other
    if (use_device() != BRM_OK) return;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                return S->stop ||
                       (!S->rebuilding && !S->queue.empty() && S->in_use + S->queue.front()->reserved <= S->budget);
            });
            if (S->stop) break;
            // FIFO: admit from the head while the budget allows
            while (!S->queue.empty() && (int)batch.size() < S->max_batch &&
                   S->in_use + S->queue.front()->reserved <= S->budget) {
                auto r = S->queue.front();
                S->queue.pop_front();
                S->in_use += r->reserved;
                batch.push_back(r);
            }
            S->inflight += 1;
            const double t = now_s();
            for (auto& r : batch) r->t_start = t;
        }
        const int st = run_batch(S, batch, s);
        std::string err = st == BRM_OK ? std::string() : std::string(brm_last_error());
        if (st != BRM_OK) (void)hipStreamSynchronize(s);
        finish_batch(S, batch, st, err);
        dev_flush_deferred();
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

// Appends B (device arrays of the caller) to the stack.
int stack_append(brm_server* S, const int* brpt, const int* bcol, const double* bval, int64_t nnzB, int K, int N,
                 hipStream_t s) {
    std::vector<brm_pair> parts;
    if (!S->bs.empty())
        parts.push_back(brm_pair{nullptr, nullptr, nullptr, 0, S->stack.brpt, S->stack.bcol, S->stack.bval,
                                 S->stack.nnzB, 0, S->stack.K, S->stack.N});
    parts.push_back(brm_pair{nullptr, nullptr, nullptr, 0, brpt, bcol, bval, nnzB, 0, K, N});
    FusedOut fo;
    BRMC(fuse_pairs(parts.data(), (int)parts.size(), nullptr, s, &fo));
    HIPC(hipStreamSynchronize(s));
    // Readers of bs and stack hold mu (workers are idle while a B is added).
    std::lock_guard<std::mutex> lk(S->mu);
    ResidentB rb{S->bs.empty() ? 0 : (int64_t)S->stack.K, nnzB, K, N};
    for (void* p : S->stack_bufs) dev_free(p);
    S->stack_bufs = fo.sc.bufs;
    fo.sc.bufs.clear();
    S->stack = fo.in;
    S->bs.push_back(rb);
    return BRM_OK;
}

}  // namespace

extern "C" {

int brm_server_create(int64_t budget_bytes, int max_batch, int streams, brm_server** out) {
    if (!out || budget_bytes <= 0 || max_batch < 1 || streams < 1 || streams > 8) {
        set_err("invalid server configuration");
        return BRM_E_INVALID;
    }
    *out = nullptr;
    BRMC(use_device());
    brm_server* S = new (std::nothrow) brm_server();
    if (!S) {
        set_err("host allocation failed");
        return BRM_E_NOMEM;
    }
    S->device = thread_device();
    S->ws_share = std::max<int64_t>(budget_bytes / 8 / streams, 1);
    S->budget = budget_bytes;
    S->max_batch = max_batch;
    if (hipStreamCreateWithFlags(&S->aux, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        set_err("hipStreamCreate failed");
        return BRM_E_DEVICE;
    }
    for (int i = 0; i < streams; ++i) S->workers.emplace_back(worker_main, S);
    *out = S;
    return BRM_OK;
}

int brm_server_add_b(brm_server* S, const int* brpt, const int* bcol, const double* bval, int64_t nnzB, int K,
                     int N, int* b_id) {
    if (!S || !b_id || K < 0 || N < 0 || nnzB < 0 || (K > 0 && !brpt) || (nnzB > 0 && (!bcol || !bval))) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    std::lock_guard<std::mutex> ak(S->aux_mu);
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->queue.empty() && S->inflight == 0; });
        S->rebuilding = true;
    }
    // validate B as a standalone operand (row pointers, column range, sorted rows)
    int st = BRM_OK;
    {
        Scratch sc;
        int* derr = nullptr;
        st = sc.alloc(&derr, 1, "error flag");
        if (st == BRM_OK) st = hip_status(launch_fill_i32(derr, 1, 0, S->aux), "fill");
        RowArgs a{};
        a.arpt = brpt;  // checked as a K x N operand through the A-side checks
        a.acol = bcol;
        a.M = K;
        a.K = N;
        a.brpt = brpt;
        if (st == BRM_OK) st = hip_status(launch_validate(a, nnzB, -1, derr, S->aux), "validate");
        if (st == BRM_OK) st = hip_status(launch_validate_sorted(brpt, bcol, K, nnzB, derr, S->aux), "validate");
        if (st == BRM_OK) st = check_dev_err(derr, S->aux, "resident B");
        if (st == BRM_OK) st = stack_append(S, brpt, bcol, bval, nnzB, K, N, S->aux);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->rebuilding = false;
    if (st == BRM_OK) *b_id = (int)S->bs.size() - 1;
    S->cv_work.notify_all();
    dev_flush_deferred();
    return st;
}

int brm_server_submit(brm_server* S, int b_id, const int* arpt, const int* acol, const double* aval, int64_t nnzA,
                      int M, int64_t* req_id) {
    const double t0 = now_s();
    if (!S || !req_id || M < 0 || nnzA < 0 || (M > 0 && !arpt) || (nnzA > 0 && (!acol || !aval)) ||
        (M == 0 && nnzA != 0)) {
        set_err("invalid argument");
        return BRM_E_INVALID;
    }
    BRMC(use_device());
    auto r = std::make_shared<Request>();
    int B_N = 0;
    {
        std::lock_guard<std::mutex> ak(S->aux_mu);
        ResidentB B;
        const int* brpt;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (b_id < 0 || b_id >= (int)S->bs.size()) {
                set_err("unknown B operand %d", b_id);
                return BRM_E_INVALID;
            }
            B = S->bs[b_id];
            brpt = S->stack.brpt + B.k0;
        }
        B_N = B.N;
        // validate A against its B and count its products (the Upper bound)
        Scratch sc;
        int* derr;
        int64_t* ipA;
        BRMC(sc.alloc(&derr, 1, "error flag"));
        BRMC(sc.alloc(&ipA, nnzA + 1, "ipA"));
        HIPC(launch_fill_i32(derr, 1, 0, S->aux));
        RowArgs a{};
        a.arpt = arpt;
        a.acol = acol;
        a.brpt = brpt;
        a.M = M;
        a.K = B.K;
        a.N = B.N;
        HIPC(launch_validate(a, nnzA, -1, derr, S->aux));
        BRMC(check_dev_err(derr, S->aux, "request"));
        HIPC(scan_alen(acol, brpt, B.K, nnzA, ipA, S->aux));
        BRMC(d2h(&r->ip, ipA + nnzA, sizeof(int64_t), S->aux));
    }
    r->b = b_id;
    r->N = B_N;
    r->arpt = arpt;
    r->acol = acol;
    r->aval = aval;
    r->nnzA = nnzA;
    r->M = M;
    r->reserved = 12 * r->ip + 8 * ((int64_t)M + 1);
    r->t_submit = t0;
    if (r->reserved > S->budget) {
        set_err("request needs %lld bytes, more than the server budget", (long long)r->reserved);
        return BRM_E_NOMEM;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    if (S->first_submit < 0) S->first_submit = t0;
    S->reqs[r->id] = r;
    S->queue.push_back(r);
    *req_id = r->id;
    S->cv_work.notify_one();
    dev_flush_deferred();
    return BRM_OK;
}

int brm_server_poll(brm_server* S, int64_t req_id, int* done) {
    if (!S || !done) return BRM_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) {
        set_err("unknown request %lld", (long long)req_id);
        return BRM_E_INVALID;
    }
    *done = it->second->done ? 1 : 0;
    return BRM_OK;
}

int brm_server_wait(brm_server* S, int64_t req_id, brm_server_result* out) {
    if (!S || !out) return BRM_E_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) {
        set_err("unknown request %lld", (long long)req_id);
        return BRM_E_INVALID;
    }
    auto r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    memset(out, 0, sizeof(*out));
    out->status = r->status;
    out->t_submit = r->t_submit;
    out->t_start = r->t_start;
    out->t_done = r->t_done;
    out->batch_size = r->batch_size;
    out->M = r->M;
    out->N = r->N;
    out->ip = r->ip;
    if (r->status != BRM_OK) {
        set_err("%s", r->err.c_str());
        return r->status;
    }
    out->crpt = r->crpt;
    out->ccol = r->ccol;
    out->cval = r->cval;
    out->nnz = r->nnz;
    return BRM_OK;
}

int brm_server_release(brm_server* S, int64_t req_id) {
    if (!S) return BRM_E_INVALID;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(req_id);
        if (it == S->reqs.end() || !it->second->done) {
            set_err("request %lld is unknown or not finished", (long long)req_id);
            return BRM_E_INVALID;
        }
        r = it->second;
        S->reqs.erase(it);
        S->in_use -= r->reserved;
        S->cv_work.notify_all();
    }
    r.reset();  // the batch's arrays are freed with the last of its results
    return BRM_OK;
}

int brm_server_stats(brm_server* S, char* buf, uint64_t cap) {
    if (!S || !buf) return BRM_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> l(S->latency);
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        const size_t k = std::min(l.size() - 1, (size_t)(q * (double)(l.size() - 1) + 0.5));
        return l[k];
    };
    const double span = S->first_submit >= 0 ? S->last_done - S->first_submit : 0.0;
    char s[768];
    snprintf(s, sizeof(s),
             "{\"completed\": %zu, \"batches\": %lld, \"mean_batch\": %.3f, \"p50_s\": %.6f, \"p99_s\": %.6f, "
             "\"max_s\": %.6f, \"span_s\": %.6f, \"total_ip\": %lld, \"in_use_bytes\": %lld, \"queued\": %zu, "
             "\"resident_b\": %zu}",
             l.size(), (long long)S->batches, S->batches ? (double)S->batched / (double)S->batches : 0.0, pct(0.50),
             pct(0.99), l.empty() ? 0.0 : l.back(), span, (long long)S->total_ip, (long long)S->in_use,
             S->queue.size(), S->bs.size());
    const size_t n = strlen(s);
    if (n + 1 > cap) return BRM_E_INVALID;
    memcpy(buf, s, n + 1);
    return BRM_OK;
}

void brm_server_destroy(brm_server* S) {
    if (!S) return;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->queue.empty() && S->inflight == 0; });
        S->stop = true;
        S->cv_work.notify_all();
    }
    for (auto& t : S->workers) t.join();
    (void)use_device();
    (void)hipDeviceSynchronize();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->reqs.clear();
    }
    delete S;
}

}  // extern "C"
