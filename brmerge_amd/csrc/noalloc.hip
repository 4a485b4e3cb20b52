// BRMerge-Precise into caller-owned outputs with no device allocation
// (brm_spgemm_precise_dev_noalloc).  No counterpart in the reference: it is
// for callers that cannot allocate device memory (graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// Every scratch buffer of the product (per-entry product offsets, row bins
// and counts, scan partials, the long-row workspace and its task tables) is
// carved from one fixed device array, `scratch`, of NOALLOC_SCRATCH_BYTES,
// defined below.  Each GPU the library runs on holds its own copy of it
// until the process ends.  In host code the array's name is only a host-side
// stand-in: its device address is taken once per device with
// hipGetSymbolAddress (at brm_init, or at the first call on a device) and
// only that address is ever passed to kernels or copies.
//
// Calls share their device's array and keep none of its contents: whatever
// must start at zero is cleared by the pipeline itself on every call.  Two
// calls on one device would race on it, so the library serialises them (one
// mutex per device); the entry returns only after its stream is idle.
#include <string.h>

#include <mutex>

#include "pipeline.h"

namespace brm {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct DeviceScratch {
    std::once_flag once;
    hipError_t err = hipSuccess;
    void* ptr = nullptr;  // device address of `scratch` on this device
    std::mutex mu;        // one call at a time on this device
};
DeviceScratch g_dev[MAX_DEVICES];

}  // namespace

// Resolves the device address of `scratch` on `device` (which must be the
// current device); once per device.
int noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) {
        set_err("device %d beyond the supported %d devices", device, MAX_DEVICES);
        return BRM_E_DEVICE;
    }
    DeviceScratch& d = g_dev[device];
    std::call_once(d.once, [&] { d.err = hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)); });
    if (d.err != hipSuccess || !d.ptr) {
        set_err("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(d.err));
        return BRM_E_DEVICE;
    }
    return BRM_OK;
}

}  // namespace brm

using namespace brm;

extern "C" {

uint64_t brm_noalloc_scratch_bytes(void) { return (uint64_t)NOALLOC_SCRATCH_BYTES; }

int brm_spgemm_precise_dev_noalloc(const int* arpt, const int* acol, const double* aval, int64_t nnzA,
                                   const int* brpt, const int* bcol, const double* bval, int64_t nnzB, int M, int K,
                                   int N, int64_t* crpt, int* ccol, double* cval, int64_t cap, int64_t* cnnz,
                                   void* stream) {
    if (!crpt || !cnnz || cap < 0 || (cap > 0 && (!ccol || !cval)) || (M > 0 && !arpt) || (K > 0 && !brpt)) {
        set_err("null pointer or negative capacity");
        return BRM_E_INVALID;
    }
    *cnnz = -1;
    BRMC(use_device());
    const int dev = thread_device();
    BRMC(noalloc_setup(dev));
    DeviceScratch& d = g_dev[dev];
    std::lock_guard<std::mutex> lk(d.mu);
    // Scratch need that is known before any kernel runs: product offsets
    // (8 bytes per A entry), per-row counts (16 bytes) and bins (4 bytes per
    // bin) of every row.
    const int64_t need = 8 * (nnzA + 1) + (16 + 4 * (BIN_COUNT - 1)) * ((int64_t)M + 1) + (int64_t)64 * 1024;
    if (need > (int64_t)NOALLOC_SCRATCH_BYTES) {
        set_err("product needs at least %lld bytes of scratch, more than the no-allocation entry's fixed %lld",
                (long long)need, (long long)NOALLOC_SCRATCH_BYTES);
        return BRM_E_NOMEM;
    }
    arena_begin(d.ptr, NOALLOC_SCRATCH_BYTES);
    DevIn in{arpt, acol, aval, nnzA, brpt, bcol, bval, nnzB, M, K, N};
    DevOut out{};
    int64_t needed = -1;
    ExtOut ext{crpt, ccol, cval, cap, &needed};
    hipStream_t s = (hipStream_t)stream;
    int st = run_spgemm(in, false, INT64_MAX, s, &out, nullptr, &ext);
    if (st != BRM_OK) {
        (void)hipStreamSynchronize(s);  // nothing of this call may still run on the scratch
        if (st == BRM_E_NOMEM && arena_remaining() >= 0 && strstr(brm_last_error(), "scratch") == nullptr) {
            set_err("the product needs more scratch memory than the no-allocation entry's fixed %lld bytes",
                    (long long)NOALLOC_SCRATCH_BYTES);
        }
    }
    arena_end();
    if (st == BRM_OK) *cnnz = out.nnz;
    else if (st == BRM_E_OVERFLOW && needed >= 0) *cnnz = needed;
    return st;
}

}  // extern "C"
