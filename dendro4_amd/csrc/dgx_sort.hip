// dgx_sort.hip -- device-wide primitives: exclusive scan and stable LSD radix
// sort of 128-bit SFC keys for gfx950.
//
// Radix sort (one-sweep): 8-bit digits, tiles of 2048 keys on 256-thread
// blocks (4 wave64s).  One up-front pass builds the global 256-bin histogram
// of every candidate digit position, and one workgroup turns them into global
// digit offsets and skips positions whose digit is the same in every key, all
// on the device.  Each remaining pass is ONE kernel: a tile counts its digits,
// publishes the counts, ranks its keys stably (8 ballots per wave, LDS
// counters across the four waves), obtains the counts of the tiles before it
// by decoupled look-back, and writes its keys run by run.
#include "dgx_common.h"

#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr int SB = 256;          // threads per block
constexpr int SITEMS = 8;        // keys per thread per tile
constexpr int STILE = SB * SITEMS;

__device__ __forceinline__ uint32_t digit_of(const Key &k, int bit) {
    return bit >= 64 ? (uint32_t)((k.hi >> (bit - 64)) & 255u) : (uint32_t)((k.lo >> bit) & 255u);
}

// ---------------------------------------------------------------- scan ----
constexpr int SCAN_B = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_B * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t &total) {
    // lds: SCAN_B/64 + 1 entries
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            uint32_t t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[blockDim.x >> 6] = acc;
    }
    __syncthreads();
    uint32_t res = x - v + lds[wid];
    total = lds[blockDim.x >> 6];
    __syncthreads();
    return res;
}

__global__ void __launch_bounds__(SCAN_B) k_scan_reduce(const uint32_t *in, size_t n, uint32_t *partial) {
    __shared__ uint32_t lds[SCAN_B / 64 + 1];
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t idx = base + (size_t)i * SCAN_B + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint32_t tot;
    block_exclusive_scan(s, lds, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single block: exclusive scan of the per-tile partials (any count).
__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t *partial, size_t np, uint32_t *d_total) {
    __shared__ uint32_t lds[1024 / 64 + 1];
    uint32_t carry = 0;
    for (size_t base = 0; base < np; base += 1024) {
        size_t idx = base + threadIdx.x;
        uint32_t v = idx < np ? partial[idx] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, lds, tot);
        if (idx < np) partial[idx] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

// The tile is staged through LDS (padded against bank conflicts), so global
// loads and stores are coalesced while each thread scans 16 consecutive items.
__device__ __forceinline__ uint32_t scan_pad(uint32_t x) { return x + (x >> 5); }
__global__ void __launch_bounds__(SCAN_B) k_scan_down(const uint32_t *in, uint32_t *out, size_t n,
                                                     const uint32_t *partial) {
    __shared__ uint32_t lds[SCAN_B / 64 + 1];
    __shared__ uint32_t t[SCAN_TILE + SCAN_TILE / 32];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const uint32_t e = (uint32_t)r * SCAN_B + tid;
        t[scan_pad(e)] = base + e < n ? in[base + e] : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = t[scan_pad(tid * SCAN_ITEMS + i)];
        s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, lds, tot) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        t[scan_pad(tid * SCAN_ITEMS + i)] = ex;
        ex += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const uint32_t e = (uint32_t)r * SCAN_B + tid;
        if (base + e < n) out[base + e] = t[scan_pad(e)];
    }
}

// One workgroup scans any n (in place allowed): 1024 threads x 16 items per
// round, carry between rounds.  Used for small arrays, where the three-kernel
// scan would be launch-bound.
constexpr int SS_B = 1024, SS_ITEMS = 16;
__global__ void __launch_bounds__(SS_B) k_scan_single(const uint32_t *in, uint32_t *out, size_t n,
                                                      uint32_t *d_total) {
    __shared__ uint32_t lds[SS_B / 64 + 1];
    uint32_t carry = 0;
    for (size_t base = 0; base < n; base += (size_t)SS_B * SS_ITEMS) {
        const size_t b0 = base + (size_t)threadIdx.x * SS_ITEMS;
        uint32_t v[SS_ITEMS];
        uint32_t t = 0;
#pragma unroll
        for (int i = 0; i < SS_ITEMS; ++i) {
            v[i] = b0 + i < n ? in[b0 + i] : 0u;
            t += v[i];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(t, lds, tot) + carry;
#pragma unroll
        for (int i = 0; i < SS_ITEMS; ++i) {
            if (b0 + i < n) out[b0 + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

// --------------------------------------------------------- radix sort ----
// Bitonic sort of at most SMALL_N keys in one workgroup (LDS), padded with
// all-ones keys.  Keys only: equal keys are identical, so stability is moot.
constexpr int SMALL_N = 2048;
__global__ void __launch_bounds__(1024) k_sort_small(Key *keys, uint32_t n) {
    __shared__ Key sk[SMALL_N];
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += 1024) sk[i] = i < n ? keys[i] : Key{~0ull, ~0ull};
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += 1024) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const Key a = sk[i], b = sk[l];
                    if (up ? key_lt(b, a) : key_lt(a, b)) {
                        sk[i] = b;
                        sk[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += 1024) keys[i] = sk[i];
}

// Global histograms of the digit positions in pmask (bit p: the byte at bit
// 8p), into 16 x 256 counters.  Four keys per thread are loaded ahead.  The
// counters have one copy per XCD (workgroup b adds into copy b % 8, the XCD
// it runs on), so the flushes of the workgroups spread over 8 places.
constexpr int GH_U = 4, GH_COPIES = 8;
__global__ void __launch_bounds__(SB) k_global_hist(const Key *keys, size_t n, uint32_t *ghist, uint32_t pmask) {
    __shared__ uint32_t h[16 * 256];
    for (int i = threadIdx.x; i < 16 * 256; i += SB) h[i] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * SB * GH_U;
    const size_t nround = (n + stride - 1) / stride * stride;  // whole waves stay in the loop
    for (size_t i0 = (size_t)blockIdx.x * SB * GH_U + threadIdx.x; i0 < nround; i0 += stride) {
        Key k[GH_U];
#pragma unroll
        for (int u = 0; u < GH_U; ++u) {
            const size_t i = i0 + (size_t)u * SB;
            k[u] = i < n ? keys[i] : Key{0ull, 0ull};
        }
#pragma unroll
        for (int u = 0; u < GH_U; ++u) {
            const bool valid = i0 + (size_t)u * SB < n;
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                if (pmask & (1u << p)) wave_hist_add(&h[p * 256], (uint32_t)((k[u].lo >> (8 * p)) & 255u), valid);
                if (pmask & (1u << (p + 8)))
                    wave_hist_add(&h[(p + 8) * 256], (uint32_t)((k[u].hi >> (8 * p)) & 255u), valid);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 16 * 256; i += SB)
        if (h[i]) atomicAdd(&ghist[(blockIdx.x % GH_COPIES) * (16 * 256) + i], h[i]);
}

// Look-back status of one (tile, digit) of one pass: bits 0-31 the count,
// bits 32-61 the pass's epoch (its index + 1; the words are zeroed once per
// sort), bits 62-63 what the count is: the tile's own count (AGG) or the
// count of that digit in all tiles up to and including it (INCL).
constexpr uint64_t ST_AGG = 1ull << 62, ST_INCL = 2ull << 62;
constexpr uint32_t SPIN_LIMIT = 1u << 20;  // polls of one status word before giving up

__device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Candidate passes of one sort and what the plan kernel derives from the
// global histograms: whether the pass's digit varies at all, which buffer
// holds its input (0 = caller's, 1 = temporary), and the exclusive global
// offset of every digit value.
struct SortPlan {
    uint32_t active[16];
    uint32_t src[17];  // src[npass]: the buffer holding the sorted result
    uint32_t gofs[16][256];
};

// One workgroup, one wave per candidate pass j (digit position
// pos[j] = (packed >> 4j) & 15): a digit position is skipped when one digit
// value holds all n keys; then thread 0 assigns the buffers in pass order.
__global__ void __launch_bounds__(1024) k_sort_plan(const uint32_t *ghist, uint32_t n, uint64_t packed, int npass,
                                                    SortPlan *plan) {
    __shared__ uint32_t act[16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (w < npass) {
        const uint32_t p = (uint32_t)(packed >> (4 * w)) & 15u;
        uint32_t c[4], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            c[q] = 0;
#pragma unroll
            for (int x = 0; x < GH_COPIES; ++x) c[q] += ghist[x * (16 * 256) + p * 256 + lane * 4 + q];
            sum += c[q];
        }
        const bool trivial = __ballot(c[0] == n || c[1] == n || c[2] == n || c[3] == n) != 0;
        uint32_t total;
        uint32_t ex = wave_excl_sum(sum, total);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            plan->gofs[w][lane * 4 + q] = ex;
            ex += c[q];
        }
        if (lane == 0) act[w] = trivial ? 0u : 1u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t src = 0;
        for (int j = 0; j < npass; ++j) {
            plan->active[j] = act[j];
            plan->src[j] = src;
            src ^= act[j];
        }
        plan->src[npass] = src;
    }
}

// One pass of the one-sweep LSD sort over the 8-bit digit at `bit`, on tiles
// of NT x SITEMS keys.  Tiles are taken in the order of a global counter, so a
// tile only ever waits for tiles that are already running.  A tile: loads its
// keys, wave w the contiguous 64 x SITEMS keys w of the tile; ranks them
// stably within the wave without workgroup barriers (per round of 64 keys: 8
// ballots give each lane its peers of equal digit, the wave's own LDS counter
// row gives the count from its earlier rounds, and the group's first lane
// advances it); turns the waves' digit counts into per-wave offsets and the
// tile's digit counts; publishes the counts (AGG); stages the keys in digit
// order in LDS; sums the counts of the tiles before it by walking back until
// a tile that already published its inclusive count (INCL); publishes its
// own inclusive count; and writes the staged keys run by run to global digit
// offset + earlier tiles' count + position in the run.  The first 256
// threads own one digit each for the counters and the look-back.
template <bool HAS_VALS, int NT>
__global__ void __launch_bounds__(NT) k_onesweep(Key *ka, Key *kb, uint32_t *va, uint32_t *vb, uint32_t n, int bit,
                                                 int j, const SortPlan *plan, uint64_t *status, uint32_t *tile_ctr,
                                                 uint32_t *fault) {
    constexpr int NW = NT / 64, TILE = NT * SITEMS;
    __shared__ uint32_t run_off[256];   // output position of staged index 0 of digit d's run, minus its start
    __shared__ uint32_t tile_off[256];  // start of digit d in the staged tile
    __shared__ uint32_t wcnt[NW][256];  // keys of digit d in wave w; then the waves' exclusive prefix
    __shared__ uint32_t scan_tmp[NW + 1];
    __shared__ uint32_t s_tile;
    __shared__ Key sk[TILE];
    __shared__ uint32_t sv[HAS_VALS ? TILE : 1];
    if (!plan->active[j]) return;  // the digit is the same in every key
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool from_b = plan->src[j] != 0;
    const Key *kin = from_b ? kb : ka;
    Key *kout = from_b ? ka : kb;
    const uint32_t *vin = from_b ? vb : va;
    uint32_t *vout = from_b ? va : vb;
    const uint64_t epoch = (uint64_t)(j + 1) << 32;
    if (tid == 0) s_tile = atomicAdd(&tile_ctr[j], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * (uint32_t)TILE, wbase = base + (uint32_t)wid * (64u * SITEMS);
    Key k[SITEMS];
    uint32_t v[SITEMS], d[SITEMS], prank[SITEMS];
#pragma unroll
    for (int r = 0; r < SITEMS; ++r) {
        const uint32_t idx = wbase + (uint32_t)(r * 64 + lane);
        const bool valid = idx < n;
        k[r] = valid ? kin[idx] : Key{0ull, 0ull};
        v[r] = (HAS_VALS && valid) ? vin[idx] : 0u;
        d[r] = digit_of(k[r], bit);
    }
    // this wave's counter row: LDS operations of one wave complete in order,
    // so the row needs no barrier between its rounds
    uint32_t *wc = wcnt[wid];
    for (int i = lane; i < 256; i += 64) wc[i] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < SITEMS; ++r) {
        const bool valid = wbase + (uint32_t)(r * 64 + lane) < n;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot((d[r] >> b) & 1u);
            peers &= ((d[r] >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        const uint32_t before = wc[d[r]];
        prank[r] = before + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wc[d[r]] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // digit tid: the waves' counts become exclusive offsets, their sum the tile's count
    uint32_t cnt = 0;
    if (tid < 256) {
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = wcnt[w][tid];
            wcnt[w][tid] = cnt;
            cnt += c;
        }
    }
    uint64_t *mine = status + (size_t)tile * 256 + (tid & 255);
    if (tid < 256) status_store(mine, (tile == 0 ? ST_INCL : ST_AGG) | epoch | cnt);
    uint32_t total;
    const uint32_t toff = block_exclusive_scan(cnt, scan_tmp, total);
    if (tid < 256) tile_off[tid] = toff;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SITEMS; ++r)
        if (wbase + (uint32_t)(r * 64 + lane) < n) {
            const uint32_t pos = tile_off[d[r]] + wc[d[r]] + prank[r];
            sk[pos] = k[r];
            if (HAS_VALS) sv[pos] = v[r];
        }
    // count of digit tid in the tiles before this one
    if (tid < 256) {
        uint32_t excl = 0;
        if (tile > 0) {
            uint32_t t = tile - 1;
            for (;;) {
                uint64_t w = status_load(status + (size_t)t * 256 + tid);
                uint32_t spins = 0;
                while ((w & ~(3ull << 62) & ~0xffffffffull) != epoch) {
                    if (++spins > SPIN_LIMIT) {
                        __hip_atomic_store(fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    w = status_load(status + (size_t)t * 256 + tid);
                }
                excl += (uint32_t)w;
                if ((w & ST_INCL) || t == 0 || spins > SPIN_LIMIT) break;
                --t;
            }
            status_store(mine, ST_INCL | epoch | (excl + cnt));
        }
        run_off[tid] = plan->gofs[j][tid] + excl - toff;
    }
    __syncthreads();
    const uint32_t tn = min((uint32_t)TILE, n - base);
    for (uint32_t i = tid; i < tn; i += NT) {
        const Key key = sk[i];
        const uint32_t pos = run_off[digit_of(key, bit)] + i;
        if (pos < n) {  // always true unless a look-back wait gave up (then the fault word is set)
            kout[pos] = key;
            if (HAS_VALS) vout[pos] = sv[i];
        }
    }
}

// Copies the result back into the caller's buffers when the last active pass
// wrote the temporary ones.
__global__ void k_sort_finish(Key *ka, const Key *kb, uint32_t *va, const uint32_t *vb, size_t n, int npass,
                              const SortPlan *plan) {
    if (!plan->src[npass]) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        ka[i] = kb[i];
        if (va) va[i] = vb[i];
    }
}

__global__ void k_set_value_u32(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace

void set_u32_async(uint32_t *p, uint32_t v, hipStream_t s) {
    k_set_value_u32<<<1, 64, 0, s>>>(p, v);
    HIP_OK(hipGetLastError());
}

void scan_exclusive_u32_async(const uint32_t *in, uint32_t *out, size_t n, uint32_t *d_total, hipStream_t s) {
    if (n == 0) {
        if (d_total) HIP_OK(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
        return;
    }
    if (n <= (size_t)SS_B * SS_ITEMS) {  // one round of the single-workgroup scan
        k_scan_single<<<1, SS_B, 0, s>>>(in, out, n, d_total);
        HIP_OK(hipGetLastError());
        return;
    }
    size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf<uint32_t> partial(nb, s);
    k_scan_reduce<<<(unsigned)nb, SCAN_B, 0, s>>>(in, n, partial.p);
    k_scan_partials<<<1, 1024, 0, s>>>(partial.p, nb, d_total);
    k_scan_down<<<(unsigned)nb, SCAN_B, 0, s>>>(in, out, n, partial.p);
    HIP_OK(hipGetLastError());
}

uint64_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s) {
    DevBuf<uint32_t> tot(1, s);
    scan_exclusive_u32_async(in, out, n, tot.p, s);
    uint32_t h = 0;
    dgx_read_back(&h, tot.p, sizeof(uint32_t), s);
    return h;
}

namespace {
void run_passes(Key *keys, uint32_t *vals, size_t n, const std::vector<int> &passes, hipStream_t s);

// Pinned host word that the sort kernels set when a look-back wait exceeds its
// bound (it never should: every wait is for a tile that is already running).
uint32_t *g_fault = nullptr;
std::once_flag g_fault_once;
uint32_t *fault_word() {
    std::call_once(g_fault_once, [] {
        void *p = nullptr;
        HIP_OK(hipHostMalloc(&p, 64, hipHostMallocMapped));
        std::memset(p, 0, 64);
        g_fault = (uint32_t *)p;
    });
    return g_fault;
}
}  // namespace

uint32_t *dgx_fault_word() { return fault_word(); }

void sort_fault_check() {
    if (g_fault && __atomic_exchange_n(g_fault, 0u, __ATOMIC_ACQ_REL)) {
        dgx_set_error("a look-back wait (radix sort or single-pass scan) exceeded its bound; that result is invalid");
        throw DgxError();
    }
}

static void check_sort_size(size_t n) {
    if (n > 0xffffffffull - 1024 * SITEMS) {  // 32-bit indices, up to the largest tile past the end
        dgx_set_error("radix_sort_keys: %zu keys exceed the 32-bit index range", n);
        throw DgxError();
    }
}

void radix_sort_keys(Key *keys, uint32_t *vals, size_t n, hipStream_t s, int begin_bit, int end_bit) {
    if (n < 2) return;
    check_sort_size(n);
    if (!vals && n <= (size_t)SMALL_N) {
        k_sort_small<<<1, 1024, 0, s>>>(keys, (uint32_t)n);
        HIP_OK(hipGetLastError());
        return;
    }
    std::vector<int> passes;
    for (int p = 0; p < 16; ++p)
        if (!(8 * p + 8 <= begin_bit || 8 * p >= end_bit)) passes.push_back(8 * p);
    run_passes(keys, vals, n, passes, s);
}


void radix_sort_pairs(Key *keys, uint32_t *vals, size_t n, hipStream_t s, uint64_t lo_max, uint64_t hi_max) {
    if (n < 2) return;
    check_sort_size(n);
    if (!vals && n <= (size_t)SMALL_N) {
        k_sort_small<<<1, 1024, 0, s>>>(keys, (uint32_t)n);
        HIP_OK(hipGetLastError());
        return;
    }
    std::vector<int> passes;
    for (int p = 0; p < 8 && (lo_max >> (8 * p)); ++p) passes.push_back(8 * p);
    for (int p = 0; p < 8 && (hi_max >> (8 * p)); ++p) passes.push_back(64 + 8 * p);
    run_passes(keys, vals, n, passes, s);
}

namespace {

// Threads per tile: larger tiles mean fewer look-back steps and fewer
// barriers per key, smaller ones more workgroups for small sorts.  A payload
// needs 4 more bytes of LDS per key, so payload sorts keep 256.
int sort_threads(size_t n, bool has_vals) {
    if (has_vals) return 256;
    if (n >= (size_t(1) << 20)) return 1024;
    if (n >= (size_t(1) << 18)) return 512;
    return 256;
}

// Workspace of one sort, zeroed by one fill: GH_COPIES x 16 x 256 global
// counters, 16 tile counters, then the plan and the look-back status words.
void run_passes(Key *keys, uint32_t *vals, size_t n, const std::vector<int> &passes, hipStream_t s) {
    if (passes.empty()) return;
    const int np = (int)passes.size();
    const int nt = sort_threads(n, vals != nullptr);
    const size_t ntiles = (n + (size_t)nt * SITEMS - 1) / ((size_t)nt * SITEMS);
    const size_t off_plan = GH_COPIES * 16 * 256 * 4 + 16 * 4;
    const size_t off_status = (off_plan + sizeof(SortPlan) + 255) / 256 * 256;
    const size_t ws_bytes = off_status + ntiles * 256 * sizeof(uint64_t);
    DevBuf<uint8_t> ws(ws_bytes, s);
    uint32_t *ghist = (uint32_t *)ws.p, *ctr = ghist + GH_COPIES * 16 * 256;
    SortPlan *plan = (SortPlan *)(ws.p + off_plan);
    uint64_t *status = (uint64_t *)(ws.p + off_status);
    HIP_OK(hipMemsetAsync(ws.p, 0, ws_bytes, s));
    uint32_t pmask = 0;
    uint64_t packed = 0;
    for (int j = 0; j < np; ++j) {
        pmask |= 1u << (passes[j] / 8);
        packed |= (uint64_t)(passes[j] / 8) << (4 * j);
    }
    const unsigned gb = (unsigned)std::min<size_t>(grid_for(n, SB * GH_U), 1024);
    k_global_hist<<<gb, SB, 0, s>>>(keys, n, ghist, pmask);
    k_sort_plan<<<1, 1024, 0, s>>>(ghist, (uint32_t)n, packed, np, plan);
    HIP_OK(hipGetLastError());
    DevBuf<Key> ktmp(n, s);
    DevBuf<uint32_t> vtmp(vals ? n : 0, s);
    uint32_t *fault = fault_word();
    const unsigned grid = (unsigned)ntiles;
    for (int j = 0; j < np; ++j) {
        auto *kp = keys, *kt = ktmp.p;
        if (vals)
            k_onesweep<true, 256><<<grid, 256, 0, s>>>(kp, kt, vals, vtmp.p, (uint32_t)n, passes[j], j, plan, status,
                                                       ctr, fault);
        else if (nt == 1024)
            k_onesweep<false, 1024><<<grid, 1024, 0, s>>>(kp, kt, nullptr, nullptr, (uint32_t)n, passes[j], j, plan,
                                                          status, ctr, fault);
        else if (nt == 512)
            k_onesweep<false, 512><<<grid, 512, 0, s>>>(kp, kt, nullptr, nullptr, (uint32_t)n, passes[j], j, plan,
                                                        status, ctr, fault);
        else
            k_onesweep<false, 256><<<grid, 256, 0, s>>>(kp, kt, nullptr, nullptr, (uint32_t)n, passes[j], j, plan,
                                                        status, ctr, fault);
        HIP_OK(hipGetLastError());
    }
    k_sort_finish<<<(unsigned)std::min<size_t>(grid_for(n, 256), 512), 256, 0, s>>>(keys, ktmp.p, vals, vtmp.p, n,
                                                                                     np, plan);
    HIP_OK(hipGetLastError());
}

}  // namespace
