// Caching device allocator.  hipMalloc/hipFree of the multi-GB outputs and
// workspaces of a large SpGEMM cost milliseconds each; blocks are therefore
// kept after release and reused by later calls of the same or smaller size.
// On allocation failure cached blocks are returned to the driver, largest
// first, until the request fits.  All library calls synchronise their stream before
// returning, so a released block is never still in use by the device.
#include <stdio.h>
#include <string.h>

#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "brm_internal.h"

namespace brm {
namespace {

std::mutex g_mu;
std::multimap<size_t, void*> g_free;         // size -> block
std::unordered_map<void*, size_t> g_used;    // block -> size
size_t g_cached = 0;
int64_t g_fail_after = -1;

size_t round_size(size_t b) {
    if (b < 256) b = 256;
    if (b < (1u << 20)) return (b + 511) & ~size_t(511);
    const size_t g = size_t(2) << 20;
    return (b + g - 1) / g * g;
}

void release_cached_locked() {
    for (auto& kv : g_free) (void)hipFree(kv.second);
    g_free.clear();
    g_cached = 0;
}

// Growable arrays released by their owners stay reserved and mapped for
// reuse (below); these give their memory back to the driver.  g_mu held.
bool vm_release_cached_one_locked();
void vm_release_cached_all_locked();
size_t vm_cached_bytes_locked();

}  // namespace

namespace {
// Arena of the calling thread (brm_spgemm_precise_dev_noalloc): while set,
// every allocation of the thread's library calls is carved from it, and
// nothing is released until the arena is cleared.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
};
thread_local Arena t_arena;
}  // namespace

void arena_begin(void* base, size_t bytes) { t_arena = Arena{(char*)base, bytes, 0}; }
void arena_end() { t_arena = Arena{}; }
int64_t arena_remaining() { return t_arena.base ? (int64_t)(t_arena.size - t_arena.used) : -1; }
size_t arena_used() { return t_arena.used; }

void* dev_alloc(size_t bytes) {
    if (t_arena.base) {
        const size_t off = (t_arena.used + 255) & ~size_t(255);
        const size_t sz = bytes > 0 ? bytes : 1;
        if (off > t_arena.size || sz > t_arena.size - off) return nullptr;
        t_arena.used = off + sz;
        return t_arena.base + off;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_fail_after >= 0) {
        if (g_fail_after == 0) { g_fail_after = -1; return nullptr; }
        g_fail_after--;
    }
    size_t sz = round_size(bytes);
    // Reuse the smallest cached block that is not more than 25% larger.
    auto it = g_free.lower_bound(sz);
    if (it != g_free.end() && it->first <= sz + sz / 4) {
        void* p = it->second;
        size_t bsz = it->first;
        g_free.erase(it);
        g_cached -= bsz;
        g_used[p] = bsz;
        return p;
    }
    // On failure, return cached blocks to the driver, largest first, only
    // until the request fits: blocks that later calls of the same sizes will
    // reuse stay cached.
    void* p = nullptr;
    while (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        if (g_free.empty()) {
            if (!vm_release_cached_one_locked()) return nullptr;
            continue;
        }
        auto big = std::prev(g_free.end());
        (void)hipFree(big->second);
        g_cached -= big->first;
        g_free.erase(big);
    }
    g_used[p] = sz;
    return p;
}

static bool vm_free(void* p);  // g_mu held

void dev_free(void* p) {
    if (!p) return;
    if (t_arena.base && (char*)p >= t_arena.base && (char*)p < t_arena.base + t_arena.size) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (vm_free(p)) return;
    auto it = g_used.find(p);
    if (it == g_used.end()) return;  // not ours
    g_free.emplace(it->second, p);
    g_cached += it->second;
    g_used.erase(it);
}

namespace {
thread_local std::vector<void*> t_deferred;
}

void dev_defer_free(void* p) {
    if (p) t_deferred.push_back(p);
}

void dev_flush_deferred() {
    for (void* p : t_deferred) dev_free(p);
    t_deferred.clear();
}

void dev_trim() {
    std::lock_guard<std::mutex> lk(g_mu);
    release_cached_locked();
    vm_release_cached_all_locked();
}

size_t dev_cached_bytes() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cached + vm_cached_bytes_locked();
}

// ---------------------------------------------------------------- growable arrays
// A device array whose final size is known only after its first elements are
// written: a virtual address range reserved for the largest size it may
// reach, backed by physical memory (hipMemCreate) only as it grows, in pieces
// of exactly VM_STEP (the reservation is a whole number of pieces; on MI355X
// access could not be granted to a larger piece).  BRMerge-Upper's row waves
// compact into C this way.  Creating and mapping the pieces costs about
// 20 ms per GiB, so a released array stays reserved and mapped in a cache,
// like the caching allocator's blocks: a later reservation of at most its
// size (and at least half of it) takes it over with its mapped pieces.  When
// memory runs out, cached arrays go back to the driver after the cached
// blocks.  All of it is under g_mu.
namespace {
constexpr size_t VM_STEP = size_t(2) << 30;
struct VmBlock {
    size_t reserved = 0;
    size_t mapped = 0;
    size_t gran = 0;  // piece size
    int dev = 0;
    std::vector<std::pair<hipMemGenericAllocationHandle_t, size_t>> handles;
};
std::unordered_map<void*, VmBlock> v_blocks;  // in use, by base address
std::vector<std::pair<void*, VmBlock>> v_cache;  // released, still mapped

hipMemAllocationProp vm_prop(int dev) {
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = dev;
    return prop;
}

void vm_release_block(void* va, VmBlock& b) {
    size_t off = 0;
    for (auto& h : b.handles) {
        (void)hipMemUnmap((char*)va + off, h.second);
        (void)hipMemRelease(h.first);
        off += h.second;
    }
    (void)hipMemAddressFree(va, b.reserved);
    (void)hipGetLastError();
}

// the cached array with the most mapped memory
bool vm_release_cached_one_locked() {
    if (v_cache.empty()) return false;
    size_t bi = 0;
    for (size_t i = 1; i < v_cache.size(); ++i)
        if (v_cache[i].second.mapped > v_cache[bi].second.mapped) bi = i;
    vm_release_block(v_cache[bi].first, v_cache[bi].second);
    v_cache.erase(v_cache.begin() + (std::ptrdiff_t)bi);
    return true;
}

void vm_release_cached_all_locked() {
    while (vm_release_cached_one_locked()) {
    }
}

size_t vm_cached_bytes_locked() {
    size_t n = 0;
    for (const auto& c : v_cache) n += c.second.mapped;
    return n;
}
}  // namespace

void* vm_reserve(size_t bytes, int dev) {
    hipMemAllocationProp prop = vm_prop(dev);
    size_t gran = 0;
    if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || gran == 0) {
        (void)hipGetLastError();
        return nullptr;
    }
    const size_t piece = (VM_STEP + gran - 1) / gran * gran;
    const size_t res = (std::max<size_t>(bytes, 1) + piece - 1) / piece * piece;
    std::lock_guard<std::mutex> lk(g_mu);
    // a cached array of this device reserving at least res and at most 2 res
    // (the smallest such)
    size_t best = v_cache.size();
    for (size_t i = 0; i < v_cache.size(); ++i) {
        const VmBlock& c = v_cache[i].second;
        if (c.dev == dev && c.gran == piece && c.reserved >= res && c.reserved <= 2 * res &&
            (best == v_cache.size() || c.reserved < v_cache[best].second.reserved))
            best = i;
    }
    if (best < v_cache.size()) {
        void* va = v_cache[best].first;
        v_blocks[va] = std::move(v_cache[best].second);
        v_cache.erase(v_cache.begin() + (std::ptrdiff_t)best);
        return va;
    }
    void* va = nullptr;
    while (hipMemAddressReserve(&va, res, 0, nullptr, 0) != hipSuccess || !va) {
        (void)hipGetLastError();
        va = nullptr;
        if (!vm_release_cached_one_locked()) return nullptr;
    }
    VmBlock b;
    b.reserved = res;
    b.gran = piece;
    b.dev = dev;
    v_blocks[va] = std::move(b);
    return va;
}

bool vm_grow(void* va, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = v_blocks.find(va);
    if (it == v_blocks.end()) return false;
    VmBlock& b = it->second;
    if (bytes > b.reserved) return false;
    const hipMemAllocationProp prop = vm_prop(b.dev);
    while (b.mapped < bytes) {
        const size_t sz = b.gran;  // one piece (the reservation is a whole number of them)
        hipMemGenericAllocationHandle_t h;
        // physical memory beside the caching allocator: when it runs out,
        // cached blocks go back to the driver, largest first, then cached
        // growable arrays
        while (hipMemCreate(&h, sz, &prop, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (!g_free.empty()) {
                auto big = std::prev(g_free.end());
                (void)hipFree(big->second);
                g_cached -= big->first;
                g_free.erase(big);
            } else if (!vm_release_cached_one_locked()) {
                return false;
            }
        }
        if (hipMemMap((char*)va + b.mapped, sz, 0, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemRelease(h);
            return false;
        }
        hipMemAccessDesc d{};
        d.location = prop.location;
        d.flags = hipMemAccessFlagsProtReadWrite;
        if (hipMemSetAccess((char*)va + b.mapped, sz, &d, 1) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemUnmap((char*)va + b.mapped, sz);
            (void)hipMemRelease(h);
            return false;
        }
        b.handles.push_back({h, sz});
        b.mapped += sz;
    }
    return true;
}

// g_mu held
static bool vm_free(void* p) {
    auto it = v_blocks.find(p);
    if (it == v_blocks.end()) return false;
    v_cache.emplace_back(p, std::move(it->second));
    v_blocks.erase(it);
    return true;
}

// ---------------------------------------------------------------- profiling
namespace {
struct PendingLaunch {
    const char* name;
    hipEvent_t a, b;
};
std::mutex p_mu;
thread_local const char* t_open_name = nullptr;
thread_local hipEvent_t t_open_ev = nullptr;
std::vector<PendingLaunch> p_pending;
std::map<std::string, std::pair<int64_t, double>> p_totals;

void prof_drain_locked() {
    for (auto& pl : p_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(pl.b) == hipSuccess && hipEventElapsedTime(&ms, pl.a, pl.b) == hipSuccess) {
            auto& t = p_totals[pl.name];
            t.first += 1;
            t.second += ms;
        }
        (void)hipEventDestroy(pl.a);
        (void)hipEventDestroy(pl.b);
    }
    p_pending.clear();
}
}  // namespace

void prof_begin(const char* name, hipStream_t s) {
    if (!params().profile || t_arena.base) return;  // no events inside the no-allocation entry
    hipEvent_t a;
    if (hipEventCreate(&a) != hipSuccess) return;
    (void)hipEventRecord(a, s);
    t_open_name = name;
    t_open_ev = a;
}

void prof_end(hipStream_t s) {
    if (!t_open_ev) return;
    hipEvent_t b;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(t_open_ev);
        t_open_ev = nullptr;
        return;
    }
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(p_mu);
    p_pending.push_back({t_open_name, t_open_ev, b});
    t_open_ev = nullptr;
    if (p_pending.size() > 4096) prof_drain_locked();
}

void dev_set_fail_after(int64_t n) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_fail_after = n;
}

}  // namespace brm

extern "C" {

void brm_prof_reset(void) {
    std::lock_guard<std::mutex> lk(brm::p_mu);
    brm::prof_drain_locked();
    brm::p_totals.clear();
}

int brm_prof_report(char* buf, uint64_t cap) {
    std::lock_guard<std::mutex> lk(brm::p_mu);
    brm::prof_drain_locked();
    std::string out = "{";
    bool first = true;
    for (auto& kv : brm::p_totals) {
        char item[256];
        snprintf(item, sizeof(item), "%s\"%s\": [%lld, %.6f]", first ? "" : ", ", kv.first.c_str(),
                 (long long)kv.second.first, kv.second.second);
        out += item;
        first = false;
    }
    out += "}";
    if (!buf || out.size() + 1 > cap) return BRM_E_INVALID;
    memcpy(buf, out.c_str(), out.size() + 1);
    return BRM_OK;
}

}  // extern "C"
