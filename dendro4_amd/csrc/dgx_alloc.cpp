// dgx_alloc.cpp -- caching, stream-ordered device allocator for the library's
// temporaries.  The build and balance passes allocate and free a few hundred
// scratch buffers per call; going through hipMallocAsync / hipFreeAsync for
// each costs tens of microseconds of host time per pair, so freed blocks are
// kept here and handed out again to the next request of the same size class
// on the same stream.  Reuse is safe in stream order: a block freed on stream
// s is only reused by work queued later on s.  Blocks beyond the cache cap,
// and all cached blocks when an allocation fails, go back to the HIP pool.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <unordered_map>

#include "dgx_common.h"

namespace {

struct Cached {
    void *p;
    hipStream_t s;
};

std::mutex g_mu;
std::unordered_map<void *, std::pair<size_t, hipStream_t>> g_live;  // pointer -> (class bytes, stream)
std::multimap<size_t, Cached> g_cache;                              // class bytes -> free block
size_t g_cached_bytes = 0;
constexpr size_t kCacheCap = size_t(16) << 30;  // bytes held back from the pool at most

size_t size_class(size_t b) {
    if (b == 0) b = 1;
    if (b <= (size_t(1) << 20)) return (b + 511) & ~size_t(511);
    const size_t g = size_t(2) << 20;
    return (b + g - 1) / g * g;
}

void trim_locked() {
    for (auto &kv : g_cache) (void)hipFreeAsync(kv.second.p, kv.second.s);
    g_cache.clear();
    g_cached_bytes = 0;
}

}  // namespace

void *dgx_alloc(size_t bytes, hipStream_t s) {
    const size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        // smallest cached block of this class or up to 25% larger, same stream
        for (auto it = g_cache.lower_bound(cls); it != g_cache.end() && it->first <= cls + cls / 4; ++it) {
            if (it->second.s != s) continue;
            void *p = it->second.p;
            g_live[p] = {it->first, s};
            g_cached_bytes -= it->first;
            g_cache.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMallocAsync(&p, cls, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_mu);
            trim_locked();
        }
        (void)hipStreamSynchronize(s);
        e = hipMallocAsync(&p, cls, s);
        if (e != hipSuccess) {
            dgx_set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
            throw DgxError();
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_live[p] = {cls, s};
    return p;
}

void dgx_free(void *p, hipStream_t s) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {  // not ours
        (void)hipFreeAsync(p, s);
        return;
    }
    const size_t cls = it->second.first;
    const hipStream_t as = it->second.second;
    g_live.erase(it);
    if (g_cached_bytes + cls > kCacheCap) {
        (void)hipFreeAsync(p, as);
        return;
    }
    g_cache.emplace(cls, Cached{p, as});
    g_cached_bytes += cls;
}

void dgx_alloc_trim() {
    std::lock_guard<std::mutex> lk(g_mu);
    trim_locked();
}

size_t dgx_alloc_cached_bytes() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cached_bytes;
}

void dgx_alloc_release_stream(hipStream_t s) {
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto it = g_cache.begin(); it != g_cache.end();) {
        if (it->second.s == s) {
            (void)hipFreeAsync(it->second.p, s);
            g_cached_bytes -= it->first;
            it = g_cache.erase(it);
        } else {
            ++it;
        }
    }
    (void)hipStreamSynchronize(s);
}
