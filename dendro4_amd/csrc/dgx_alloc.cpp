// dgx_alloc.cpp -- caching, stream-ordered device allocator for the library's
// temporaries.  The build and balance passes allocate and free a few hundred
// scratch buffers per call; going through hipMallocAsync / hipFreeAsync for
// each costs tens of microseconds of host time per pair, so freed blocks are
// kept here and handed out again to the next request of the same size class
// on the same stream.  Reuse is safe in stream order: a block freed on stream
// s is only reused by work queued later on s.  When the cache is at its cap,
// blocks cached for other streams go back to the HIP pool (on their streams)
// to make room, and a block that still does not fit goes back on the stream it
// was freed on; when an allocation fails, the blocks cached for the
// allocating stream are returned and the allocation is retried.
//
// Accounting: every allocation carries the calling thread's current tag
// (dgx_alloc_set_tag); live and peak bytes are kept per tag, whichever thread
// frees a block, so a caller can measure what one piece of work holds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "dgx_common.h"

namespace {

struct Live {
    size_t bytes;
    hipStream_t s;
    uint64_t tag;
};
struct TagUse {
    int64_t live = 0, peak = 0;
};

std::mutex g_mu;
std::unordered_map<void *, Live> g_live;
std::unordered_map<uint64_t, TagUse> g_tags;  // tag 0 (untagged) is not tracked
// free blocks per stream: class bytes -> block.  Per stream, so that an
// allocation and a make-room pass look only at the blocks they can use
std::unordered_map<hipStream_t, std::multimap<size_t, void *>> g_cache;
size_t g_cached_bytes = 0;
size_t g_live_bytes = 0, g_live_peak = 0;     // bytes handed out and not freed, and their peak
// Bytes held back from the pool at most: an eighth of the device's memory
// (36 GB on MI355X), taken on first use.  Set-up allocates and frees buffers
// of up to gigabytes, and going back to the pool for them costs more than the
// work they hold: on the 2,000-request served stream an 8 GB cap served 29
// problems/s and this one 59/s.
size_t g_cap = 0;
size_t cache_cap_locked() {
    if (!g_cap) {
        size_t fr = 0, tot = 0;
        g_cap = (hipMemGetInfo(&fr, &tot) == hipSuccess && tot) ? tot / 8 : (size_t(8) << 30);
        (void)hipGetLastError();
    }
    return g_cap;
}

std::unordered_map<hipStream_t, int> g_capturing;   // streams inside a graph capture
thread_local hipStream_t t_free_override = nullptr;  // see dgx_alloc_free_stream
thread_local uint64_t t_tag = 0;                     // see dgx_alloc_set_tag

size_t size_class(size_t b) {
    if (b == 0) b = 1;
    if (b <= (size_t(1) << 20)) return (b + 511) & ~size_t(511);
    const size_t g = size_t(2) << 20;
    return (b + g - 1) / g * g;
}

void note_alloc_locked(void *p, size_t cls, hipStream_t s) {
    g_live[p] = Live{cls, s, t_tag};
    g_live_bytes += cls;
    g_live_peak = std::max(g_live_peak, g_live_bytes);
    if (t_tag) {
        TagUse &u = g_tags[t_tag];
        u.live += (int64_t)cls;
        if (u.live > u.peak) u.peak = u.live;
    }
}

// Return the blocks cached for stream s to the pool (queued on s).
void trim_stream_locked(hipStream_t s) {
    auto m = g_cache.find(s);
    if (m == g_cache.end()) return;
    for (auto &kv : m->second) {
        (void)hipFreeAsync(kv.second, s);
        g_cached_bytes -= kv.first;
    }
    g_cache.erase(m);
}

// The cap is full: return blocks cached for other streams to the pool (each
// queued on its own stream) until `need` more bytes fit.  Blocks of a stream
// that no longer allocates (a finished set-up, a closed request slot) would
// otherwise hold the cache, and the streams still working would go through
// the pool for every block.
// Streams inside a graph capture are skipped: a free queued there would join
// the capture.
// Largest blocks first, so that few frees make the room.
void make_room_locked(size_t need, hipStream_t keep) {
    for (auto m = g_cache.begin(); m != g_cache.end() && g_cached_bytes + need > cache_cap_locked();) {
        const hipStream_t st = m->first;
        if (st == keep || g_capturing.count(st)) {
            ++m;
            continue;
        }
        auto &blocks = m->second;
        while (!blocks.empty() && g_cached_bytes + need > cache_cap_locked()) {
            auto it = std::prev(blocks.end());
            (void)hipFreeAsync(it->second, st);
            g_cached_bytes -= it->first;
            blocks.erase(it);
        }
        if (blocks.empty()) m = g_cache.erase(m);
        else ++m;
    }
}

}  // namespace

void *dgx_alloc(size_t bytes, hipStream_t s) {
    const size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        // smallest cached block of this class or up to 25% larger (blocks of
        // up to 1 MiB) or twice as large (larger blocks), same stream
        auto m = g_cache.find(s);
        if (m != g_cache.end()) {
            auto it = m->second.lower_bound(cls);
            const size_t slack = cls <= (size_t(1) << 20) ? cls / 4 : cls;
            if (it != m->second.end() && it->first <= cls + slack) {
                void *p = it->second;
                note_alloc_locked(p, it->first, s);
                g_cached_bytes -= it->first;
                m->second.erase(it);
                return p;
            }
        }
    }
    void *p = nullptr;
    hipError_t e = hipMallocAsync(&p, cls, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_mu);
            trim_stream_locked(s);
        }
        (void)hipStreamSynchronize(s);
        e = hipMallocAsync(&p, cls, s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            dgx_set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
            throw DgxError();
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    note_alloc_locked(p, cls, s);
    return p;
}

void dgx_free(void *p, hipStream_t s) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {  // not ours
        (void)hipFreeAsync(p, t_free_override ? t_free_override : s);
        return;
    }
    const size_t cls = it->second.bytes;
    g_live_bytes -= cls;
    const hipStream_t as = t_free_override ? t_free_override : it->second.s;
    if (it->second.tag) {
        auto t = g_tags.find(it->second.tag);
        if (t != g_tags.end()) {
            t->second.live -= (int64_t)cls;
            if (t->second.live <= 0) g_tags.erase(t);
        }
    }
    g_live.erase(it);
    if (g_cached_bytes + cls > cache_cap_locked()) make_room_locked(cls, as);
    if (g_cached_bytes + cls > cache_cap_locked()) {
        (void)hipFreeAsync(p, as);
        return;
    }
    g_cache[as].emplace(cls, p);
    g_cached_bytes += cls;
}

// Blocks cached for a stream inside a graph capture stay: a free queued there
// would join the capture.
void dgx_alloc_trim() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto m = g_cache.begin(); m != g_cache.end();) {
        if (g_capturing.count(m->first)) {
            ++m;
            continue;
        }
        for (auto &kv : m->second) {
            (void)hipFreeAsync(kv.second, m->first);
            g_cached_bytes -= kv.first;
        }
        m = g_cache.erase(m);
    }
}

size_t dgx_alloc_cached_bytes() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_cached_bytes;
}

void dgx_alloc_live(size_t *live, size_t *peak, int reset_peak) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (live) *live = g_live_bytes;
    if (peak) *peak = g_live_peak;
    if (reset_peak) g_live_peak = g_live_bytes;
}

void dgx_alloc_release_stream(hipStream_t s) {
    (void)hipStreamSynchronize(s);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        trim_stream_locked(s);
    }
    (void)hipStreamSynchronize(s);
}

hipStream_t dgx_alloc_free_stream(hipStream_t s) {
    hipStream_t old = t_free_override;
    t_free_override = s;
    return old;
}
hipStream_t dgx_alloc_free_stream_get() { return t_free_override; }

uint64_t dgx_alloc_get_tag() { return t_tag; }

void dgx_alloc_capture_begin(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    ++g_capturing[s];
}
void dgx_alloc_capture_end(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_capturing.find(s);
    if (it != g_capturing.end() && --it->second <= 0) g_capturing.erase(it);
}

uint64_t dgx_alloc_set_tag(uint64_t tag) {
    uint64_t old = t_tag;
    t_tag = tag;
    return old;
}
int64_t dgx_alloc_tag_live(uint64_t tag) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tags.find(tag);
    return it == g_tags.end() ? 0 : it->second.live;
}
int64_t dgx_alloc_tag_peak(uint64_t tag) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tags.find(tag);
    return it == g_tags.end() ? 0 : it->second.peak;
}

// ------------------------------------------------------------- read-back ----
namespace {
constexpr size_t kReadBackBytes = 4096;
std::mutex g_rb_mu;
std::vector<void *> g_rb_free;  // pinned buffers of threads that have ended
struct ReadBackBuf {
    void *p = nullptr;
    ~ReadBackBuf() {
        if (!p) return;
        std::lock_guard<std::mutex> lk(g_rb_mu);
        g_rb_free.push_back(p);
    }
};
void *read_back_buf() {
    thread_local ReadBackBuf b;
    if (!b.p) {
        {
            std::lock_guard<std::mutex> lk(g_rb_mu);
            if (!g_rb_free.empty()) {
                b.p = g_rb_free.back();
                g_rb_free.pop_back();
            }
        }
        if (!b.p) HIP_OK(hipHostMalloc(&b.p, kReadBackBytes, hipHostMallocDefault));
    }
    return b.p;
}
}  // namespace

void dgx_read_back(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    if (bytes > kReadBackBytes) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        return;
    }
    void *buf = read_back_buf();
    HIP_OK(hipMemcpyAsync(buf, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::memcpy(dst, buf, bytes);
}
