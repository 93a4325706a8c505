// Device memory cache for handle allocations.
//
// A handle allocates its point arrays, tables, rows and scratch once per stage and frees all of them when
// it is destroyed; a caller that smooths one cloud after another would otherwise pay hipMalloc and hipFree
// (each a device-wide synchronization) for every buffer of every cloud. Blocks released by a handle are
// kept per device and handed to later requests of a similar size. Release happens only after the device
// has drained (DBuf::free synchronizes first), so a cached block is never in use by in-flight work and can
// be handed to any stream. Cached bytes per device are capped (EGT_DEVICE_CACHE_MB, default 32768); a
// failed hipMalloc returns the device's cached blocks to the driver and tries once more.
// egt_release_cached_memory() returns every cached block.
#include <cstdlib>
#include <map>
#include <mutex>
#include "egt_internal.h"

namespace egt {
namespace {

struct Cache {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void*>> free_blocks;  // device -> (capacity, block)
    std::map<int, size_t> cached;                             // device -> cached bytes
};

Cache& cache() {
    static Cache* c = new Cache;  // never destroyed: HIP may already be torn down at static destruction
    return *c;
}

size_t cache_cap_bytes() {
    static size_t cap = [] {
        const char* e = std::getenv("EGT_DEVICE_CACHE_MB");
        const long long mb = e ? std::atoll(e) : 32768;
        return (size_t)(mb < 0 ? 0 : mb) << 20;
    }();
    return cap;
}

constexpr size_t GRAIN = 256;

// Frees the cached blocks of device dev (the caller holds the lock and has dev current).
size_t drop_device(Cache& c, int dev) {
    size_t freed = 0;
    auto& fb = c.free_blocks[dev];
    for (auto& kv : fb) {
        (void)hipFree(kv.second);
        freed += kv.first;
    }
    fb.clear();
    c.cached[dev] = 0;
    return freed;
}

}  // namespace

void* device_alloc(size_t bytes, size_t* capacity) {
    bytes = (std::max<size_t>(bytes, 1) + GRAIN - 1) / GRAIN * GRAIN;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    Cache& c = cache();
    {
        std::lock_guard<std::mutex> lock(c.mu);
        auto& fb = c.free_blocks[dev];
        auto it = fb.lower_bound(bytes);
        // reuse the smallest cached block that fits, unless it would waste more than the request itself
        if (it != fb.end() && it->first <= 2 * bytes + (1u << 20)) {
            void* p = it->second;
            *capacity = it->first;
            c.cached[dev] -= it->first;
            fb.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lock(c.mu);
            drop_device(c, dev);
        }
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw HipError{e, "hipMalloc of " + std::to_string(bytes) + " bytes"};
    }
    *capacity = bytes;
    return p;
}

void device_release(void* p, size_t capacity) {
    if (!p) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    Cache& c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    if (c.cached[dev] + capacity > cache_cap_bytes()) {
        (void)hipFree(p);
        return;
    }
    c.free_blocks[dev].emplace(capacity, p);
    c.cached[dev] += capacity;
}

size_t device_cache_release_all() {
    Cache& c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    size_t freed = 0;
    for (auto& kv : c.free_blocks) {
        if (kv.second.empty()) continue;
        if (hipSetDevice(kv.first) != hipSuccess) continue;
        (void)hipDeviceSynchronize();
        freed += drop_device(c, kv.first);
    }
    (void)hipSetDevice(cur);
    return freed;
}

size_t device_cache_bytes() {
    Cache& c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    size_t b = 0;
    for (auto& kv : c.cached) b += kv.second;
    return b;
}

}  // namespace egt
