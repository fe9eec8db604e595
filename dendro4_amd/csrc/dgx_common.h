// dgx_common.h -- shared device/host definitions for the MI355X octree-FEM path.
//
// Octants are carried on the device as 128-bit space-filling-curve keys:
//   hi: digits of levels 1..21 (3 bits each, level t at bits [63-3t, 65-3t]), bit 63 = 0
//   lo: digits of levels 22..31 (level t at bits [63-3(t-21), 65-3(t-21)]), level in bits 0..4
// A digit is the rank of the child along the curve (Hilbert rank under the
// parent's rotation, or the Morton index z<<2|y<<1|x for Morton order), so
// unsigned (hi, lo) comparison of keys reproduces ot::TreeNode::operator<
// (include/TreeNode.tcc:122-238): ancestors precede descendants, and two
// disjoint octants compare by the child rank at the level below their nearest
// common ancestor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define DGX_MAX_DEPTH 31

struct Key {
    uint64_t hi, lo;
};

// SFC tables, one set per curve (0 = Hilbert, 1 = Morton).
//   rank[r*8+m]  : rank of Morton child m under rotation r
//   morton[r*8+k]: Morton index of the rank-k child under rotation r
//   crot[r*8+m]  : rotation inside Morton child m
// The Hilbert values are the reference's table data (src/hcurvedata.cpp:329-546).
struct SfcTables {
    uint8_t rank[24 * 8];
    uint8_t morton[24 * 8];
    uint8_t crot[24 * 8];
};

const SfcTables &host_sfc(int curve);

enum { CURVE_HILBERT = 0, CURVE_MORTON = 1 };

__host__ __device__ inline bool key_lt(const Key &a, const Key &b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}
__host__ __device__ inline bool key_le(const Key &a, const Key &b) { return !key_lt(b, a); }
__host__ __device__ inline bool key_eq(const Key &a, const Key &b) { return a.hi == b.hi && a.lo == b.lo; }
__host__ __device__ inline uint32_t key_level(const Key &k) { return (uint32_t)(k.lo & 31u); }

__host__ __device__ inline uint32_t key_digit(const Key &k, uint32_t t) {
    return t <= 21 ? (uint32_t)((k.hi >> (63 - 3 * t)) & 7u) : (uint32_t)((k.lo >> (63 - 3 * (t - 21))) & 7u);
}
__host__ __device__ inline void key_set_digit(Key &k, uint32_t t, uint32_t d) {
    if (t <= 21) {
        uint32_t s = 63 - 3 * t;
        k.hi = (k.hi & ~(7ull << s)) | ((uint64_t)d << s);
    } else {
        uint32_t s = 63 - 3 * (t - 21);
        k.lo = (k.lo & ~(7ull << s)) | ((uint64_t)d << s);
    }
}
// Keep digits 1..L, clear the rest, set level L.
__host__ __device__ inline Key key_truncate(const Key &k, uint32_t L) {
    Key r;
    if (L <= 21) {
        uint32_t s = 63 - 3 * L;
        r.hi = (s >= 64) ? 0 : (k.hi & ~((1ull << s) - 1ull));
        r.lo = L;
    } else {
        r.hi = k.hi;
        uint32_t s = 63 - 3 * (L - 21);
        r.lo = (k.lo & ~((1ull << s) - 1ull)) | L;
    }
    return r;
}
// Number of leading digits two keys share (levels ignored); 31 if all equal.
__host__ __device__ inline uint32_t key_common_levels(const Key &a, const Key &b) {
    uint64_t x = a.hi ^ b.hi;
    if (x) {
#ifdef __HIP_DEVICE_COMPILE__
        uint32_t p = 63 - __clzll(x);
#else
        uint32_t p = 63 - __builtin_clzll(x);
#endif
        return (65 - p) / 3 - 1;
    }
    x = (a.lo ^ b.lo) & ~31ull;
    if (!x) return 31;
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t p = 63 - __clzll(x);
#else
    uint32_t p = 63 - __builtin_clzll(x);
#endif
    return 21 + (65 - p) / 3 - 1;
}
__host__ __device__ inline bool key_is_ancestor(const Key &a, const Key &b) {
    uint32_t la = key_level(a);
    return la < key_level(b) && key_eq(key_truncate(b, la), a);
}
__host__ __device__ inline Key key_parent(const Key &k) {
    uint32_t L = key_level(k);
    return key_truncate(k, L ? L - 1 : 0);
}
__host__ __device__ inline Key key_first_child(const Key &k) {
    Key r = k;
    r.lo = (r.lo & ~31ull) | (key_level(k) + 1);
    return r;
}
// getNext in digit space (TreeNode.tcc:460-550): the SFC successor at the same
// or a coarser level.  Returns false when k holds the last cell of the domain.
__host__ __device__ inline bool key_next(const Key &k, Key &out) {
    uint32_t L = key_level(k);
    for (uint32_t t = L; t >= 1; --t) {
        uint32_t d = key_digit(k, t);
        if (d < 7) {
            Key r = key_truncate(k, t);
            key_set_digit(r, t, d + 1);
            out = r;
            return true;
        }
    }
    return false;
}

// Encode an octant given by anchor (x,y,z) and level in a domain of 2^md.
__host__ __device__ inline Key sfc_encode(const SfcTables &T, uint32_t x, uint32_t y, uint32_t z, uint32_t lev,
                                          uint32_t md) {
    Key k{0ull, 0ull};
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        uint32_t b = md - t;
        uint32_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        uint32_t r = T.rank[rot * 8 + m];
        rot = T.crot[rot * 8 + m];
        if (t <= 21) k.hi |= (uint64_t)r << (63 - 3 * t);
        else k.lo |= (uint64_t)r << (63 - 3 * (t - 21));
    }
    k.lo |= lev;
    return k;
}
__host__ __device__ inline void sfc_decode(const SfcTables &T, const Key &k, uint32_t md, uint32_t &x, uint32_t &y,
                                           uint32_t &z, uint32_t &lev) {
    lev = key_level(k);
    x = y = z = 0;
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        uint32_t r = key_digit(k, t);
        uint32_t m = T.morton[rot * 8 + r];
        uint32_t b = md - t;
        x |= (m & 1u) << b;
        y |= ((m >> 1) & 1u) << b;
        z |= ((m >> 2) & 1u) << b;
        rot = T.crot[rot * 8 + m];
    }
}
// Morton keys without tables (digit = z<<2|y<<1|x of the anchor bits).  Bit
// j of a 21-bit coordinate slice goes to bit 3j of the key word (x; y one
// above, z two above): the hi word takes anchor bits md-1 .. md-21 (levels
// 1..21), the lo word anchor bits md-22 .. 0 (levels 22..md) at bit
// 126 - 3 md + 3 b.  Anchor bits finer than the level are ignored, as are
// digits finer than the level when decoding.
__host__ __device__ inline uint64_t morton_spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
__host__ __device__ inline uint32_t morton_compact21(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}
__host__ __device__ inline Key morton_encode(uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    // anchor bits md-1 .. md-lev
    const uint64_t keep = ((1ull << md) - 1ull) & ~((1ull << (md - lev)) - 1ull);
    const uint64_t xm = x & keep, ym = y & keep, zm = z & keep;
    Key k;
    const uint64_t xh = md >= 21 ? xm >> (md - 21) : xm << (21 - md);
    const uint64_t yh = md >= 21 ? ym >> (md - 21) : ym << (21 - md);
    const uint64_t zh = md >= 21 ? zm >> (md - 21) : zm << (21 - md);
    k.hi = morton_spread21(xh) | morton_spread21(yh) << 1 | morton_spread21(zh) << 2;
    k.lo = lev;
    if (lev > 21) {
        const uint64_t low = (1ull << (md - 21)) - 1ull;
        const uint32_t sh = 126 - 3 * md;
        k.lo |= (morton_spread21(xm & low) | morton_spread21(ym & low) << 1 | morton_spread21(zm & low) << 2) << sh;
    }
    return k;
}
__host__ __device__ inline void morton_decode(const Key &k, uint32_t md, uint32_t &x, uint32_t &y, uint32_t &z,
                                              uint32_t &lev) {
    lev = key_level(k);
    const uint32_t lh = lev < 21 ? lev : 21;
    // digits 1..lh of the hi word: bits 63 - 3 lh .. 62
    const uint64_t hi = lh ? k.hi & ~((1ull << (63 - 3 * lh)) - 1ull) : 0ull;
    uint64_t xs = morton_compact21(hi), ys = morton_compact21(hi >> 1), zs = morton_compact21(hi >> 2);
    if (md >= 21) {
        xs <<= md - 21;
        ys <<= md - 21;
        zs <<= md - 21;
    } else {
        xs >>= 21 - md;
        ys >>= 21 - md;
        zs >>= 21 - md;
    }
    if (lev > 21) {
        // digits 22..lev of the lo word: bits 126 - 3 lev .. 62
        const uint64_t lo = k.lo & ~((1ull << (126 - 3 * lev)) - 1ull);
        const uint32_t sh = 126 - 3 * md;
        xs |= morton_compact21(lo >> sh);
        ys |= morton_compact21(lo >> (sh + 1));
        zs |= morton_compact21(lo >> (sh + 2));
    }
    x = (uint32_t)xs;
    y = (uint32_t)ys;
    z = (uint32_t)zs;
}

// ---------------------------------------------------------------------------
// Wave64 helpers: aggregate per-lane counts so that one lane per wave touches
// a global counter.
// ---------------------------------------------------------------------------
__device__ inline uint32_t wave_excl_sum(uint32_t v, uint32_t &total) {
    const int lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}
// Exclusive scan over the workgroup (a multiple of 64 threads); lds holds
// blockDim.x / 64 + 1 entries; every thread must call it.
__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t *lds, uint32_t &total) {
    uint32_t wt;
    const uint32_t ex = wave_excl_sum(v, wt);
    const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (__lane_id() == 0) lds[wid] = wt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; ++w) {
            const uint32_t t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[nw] = acc;
    }
    __syncthreads();
    const uint32_t r = lds[wid] + ex;
    total = lds[nw];
    __syncthreads();
    return r;
}
// Decoupled look-back over tiles for one 32-bit value per tile, for
// single-pass scans.  status[t] holds bits 0-31 a count, bits 32-61 the call's
// epoch (status words are zeroed once and every call uses a new epoch), bits
// 62-63 whether the count is the tile's own (1) or inclusive of all tiles
// before it (2).  Called by one thread per tile, with tiles numbered in the
// order they started (so every wait is for a running tile); returns the sum
// of the earlier tiles' values.  A wait past its bound sets *fault.
constexpr uint64_t LB_OWN = 1ull << 62, LB_INCL = 2ull << 62;
__device__ inline uint32_t lookback_prefix(uint64_t *status, uint32_t tile, uint32_t value, uint32_t epoch,
                                           uint32_t *fault) {
    const uint64_t ep = (uint64_t)epoch << 32;
    uint64_t *mine = status + tile;
    if (tile == 0) {
        __hip_atomic_store(mine, LB_INCL | ep | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    __hip_atomic_store(mine, LB_OWN | ep | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0, t = tile - 1;
    for (;;) {
        uint64_t w = __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        while ((w & ~(3ull << 62) & ~0xffffffffull) != ep) {
            if (++spins > (1u << 20)) {
                __hip_atomic_store(fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            w = __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        excl += (uint32_t)w;
        if ((w & LB_INCL) || t == 0 || spins > (1u << 20)) break;
        --t;
    }
    __hip_atomic_store(mine, LB_INCL | ep | (excl + value), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}
// Reserve n slots on *counter for this thread with one atomic per workgroup;
// returns the first slot.  Every thread of the workgroup must call it; lds
// holds blockDim.x / 64 + 1 entries.
__device__ inline uint32_t block_reserve(uint32_t *counter, uint32_t n, uint32_t *lds) {
    uint32_t wtot;
    const uint32_t ex = wave_excl_sum(n, wtot);
    const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (__lane_id() == 0) lds[wid] = wtot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; ++w) {
            const uint32_t t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[nw] = acc ? atomicAdd(counter, acc) : 0u;
    }
    __syncthreads();
    const uint32_t r = lds[nw] + lds[wid] + ex;
    __syncthreads();  // lds may be reused by the caller
    return r;
}

// Histogram add of one count per valid lane into h[v] (LDS or global).  The
// few most common values of the wave are counted with one atomic each
// (ballot match) while each still holds at least 8 lanes, the rest with
// per-lane atomics: clustered data do not serialise 64 lanes on one address,
// and scattered data pay no match rounds that save nothing.  Every lane of
// the wave must call it.
__device__ inline void wave_hist_add(uint32_t *h, uint32_t v, bool valid) {
    uint64_t act = __ballot(valid);
    for (int it = 0; it < 4 && act; ++it) {
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t vl = __shfl(v, leader, 64);
        const uint64_t m = __ballot(valid && v == vl) & act;
        if (__popcll(m) < 8) break;  // scattered values: one per-lane atomic is cheaper than more rounds
        if (__lane_id() == leader) atomicAdd(&h[vl], (uint32_t)__popcll(m));
        act &= ~m;
    }
    if ((act >> __lane_id()) & 1ull) atomicAdd(&h[v], 1u);
}
// Slot for this lane in bucket v: base counters cnt[] (LDS) are advanced by
// one atomic per distinct value in the wave; lanes keep their order within a
// bucket.  Every lane of the wave must call it.
__device__ inline uint32_t wave_bucket_slot(uint32_t *cnt, uint32_t v, bool valid) {
    const int lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t act = __ballot(valid);
    uint32_t slot = 0;
    while (act) {
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t vl = __shfl(v, leader, 64);
        const uint64_t m = __ballot(valid && v == vl) & act;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cnt[vl], (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        if ((m >> lane) & 1ull) slot = base + (uint32_t)__popcll(m & lt);
        act &= ~m;
    }
    return slot;
}

// ---------------------------------------------------------------------------
// Host-side error plumbing and device buffers.
// ---------------------------------------------------------------------------
void dgx_set_error(const char *fmt, ...);

#define HIP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            dgx_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));   \
            throw DgxError();                                                                          \
        }                                                                                              \
    } while (0)

struct DgxError {};

// Stream-ordered device allocation (dgx_alloc.cpp): a size-class cache in
// front of the library's HIP memory pool.
void *dgx_alloc(size_t bytes, hipStream_t s);
void dgx_free(void *p, hipStream_t s);
void dgx_alloc_trim();            // return the cached blocks to the pool (not those of capturing streams)
size_t dgx_alloc_cached_bytes();  // bytes held in the cache
// Bytes of the library's device allocations in use, and their peak since the
// process started or the last reset.
void dgx_alloc_live(size_t *live, size_t *peak, int reset_peak);
// Return the cached blocks of stream s to the pool before s is destroyed
// (no live allocation may still be tagged with s).
void dgx_alloc_release_stream(hipStream_t s);
// While set on this thread, frees are queued on / cached for stream s instead
// of the stream a block was allocated on (for objects whose work has finished
// and whose own stream may be in use by another thread, e.g. under graph
// capture).  Returns the previous setting.
hipStream_t dgx_alloc_free_stream(hipStream_t s);
hipStream_t dgx_alloc_free_stream_get();
// Accounting tags: allocations made while a thread's tag is set are counted
// under it (live and peak bytes), whichever thread frees them; tag 0 = none.
// The counters of a tag disappear when its live bytes return to zero.
uint64_t dgx_alloc_set_tag(uint64_t tag);  // returns the previous tag
uint64_t dgx_alloc_get_tag();              // the calling thread's tag
// Brackets a graph capture on s (call begin before hipStreamBeginCapture and
// end after hipStreamEndCapture): the allocator then queues no frees on s.
void dgx_alloc_capture_begin(hipStream_t s);
void dgx_alloc_capture_end(hipStream_t s);
int64_t dgx_alloc_tag_live(uint64_t tag);
int64_t dgx_alloc_tag_peak(uint64_t tag);

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t s = 0;
    DevBuf() {}
    DevBuf(size_t count, hipStream_t st) { alloc(count, st); }
    void alloc(size_t count, hipStream_t st) {
        release();
        s = st;
        n = count;
        p = count ? (T *)dgx_alloc(count * sizeof(T), st) : nullptr;
    }
    void release() {
        if (p) dgx_free(p, s);
        p = nullptr;
        n = 0;
    }
    T *detach() {
        T *q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), s(o.s) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; s = o.s; o.p = nullptr; o.n = 0; }
        return *this;
    }
};

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------
// Primitives (dgx_sort.hip)
// ---------------------------------------------------------------------------
// Stable LSD radix sort of 128-bit keys with an optional 32-bit payload.
// Bits [begin_bit, end_bit) of the 128-bit value (hi:lo) are sorted, one pass
// per 8-bit digit position that overlaps the range; positions in which every
// key agrees are detected and skipped on the device.  Nothing is read back:
// the sort is stream-ordered.  At most 2048 keys without payload are sorted in
// one workgroup.
void radix_sort_keys(Key *keys, uint32_t *vals, size_t n, hipStream_t s, int begin_bit = 0, int end_bit = 128);
// The same sort for keys that pack two bounded integers: lo <= lo_max and
// hi <= hi_max; only the bytes those bounds can reach get a pass.
void radix_sort_pairs(Key *keys, uint32_t *vals, size_t n, hipStream_t s, uint64_t lo_max, uint64_t hi_max);
// The pinned host word the look-back kernels set on a wait past its bound.
uint32_t *dgx_fault_word();
// Copies bytes from device memory to host memory on s and waits for it.  Up
// to 4 KiB go through a pinned buffer of the calling thread (kept for later
// threads when it exits), so the copy needs no staging.
void dgx_read_back(void *dst, const void *src, size_t bytes, hipStream_t s);
// Throws DgxError if a sort kernel reported a look-back wait that exceeded its
// bound since the last check (the flag is cleared).
void sort_fault_check();
// Exclusive scan of n uint32 values; returns the total (synchronises the stream).
// *p = v in stream order (a kernel, so it is safe under graph capture and
// needs no host staging).
void set_u32_async(uint32_t *p, uint32_t v, hipStream_t s);
uint64_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s);
// Same, leaving the total in device memory (no host synchronisation).
void scan_exclusive_u32_async(const uint32_t *in, uint32_t *out, size_t n, uint32_t *d_total, hipStream_t s);
