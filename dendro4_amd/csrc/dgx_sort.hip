// dgx_sort.hip -- device-wide primitives: exclusive scan and stable LSD radix
// sort of 128-bit SFC keys for gfx950.
//
// Radix sort: 8-bit digits, 256-thread blocks (4 wave64s).  One up-front pass
// builds the global 256-bin histogram of every 8-bit digit position; passes
// whose digit is constant over all keys are skipped.  Each remaining pass is
// block histogram -> digit-major exclusive scan -> stable scatter, where the
// scatter ranks keys inside a wave with 8 ballots (one per digit bit) and
// orders the four waves of a block through LDS counters.
#include "dgx_common.h"

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

__global__ void __launch_bounds__(SCAN_B) k_scan_down(const uint32_t *in, uint32_t *out, size_t n,
                                                     const uint32_t *partial) {
    __shared__ uint32_t lds[SCAN_B / 64 + 1];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, lds, tot) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t idx = base + i;
        if (idx < n) out[idx] = ex;
        ex += v[i];
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

// Global histograms of all 16 digit positions (16 x 256 counters).
__global__ void __launch_bounds__(SB) k_global_hist(const Key *keys, size_t n, uint32_t *ghist) {
    __shared__ uint32_t h[16 * 256];
    for (int i = threadIdx.x; i < 16 * 256; i += SB) h[i] = 0;
    __syncthreads();
    size_t stride = (size_t)gridDim.x * SB;
    const size_t nround = (n + stride - 1) / stride * stride;  // whole waves stay in the loop
    for (size_t i = (size_t)blockIdx.x * SB + threadIdx.x; i < nround; i += stride) {
        const bool valid = i < n;
        Key k{0ull, 0ull};
        if (valid) k = keys[i];
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            wave_hist_add(&h[p * 256], (uint32_t)((k.lo >> (8 * p)) & 255u), valid);
            wave_hist_add(&h[(p + 8) * 256], (uint32_t)((k.hi >> (8 * p)) & 255u), valid);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 16 * 256; i += SB)
        if (h[i]) atomicAdd(&ghist[i], h[i]);
}

__global__ void __launch_bounds__(SB) k_block_hist(const Key *keys, size_t n, int bit, uint32_t *bhist,
                                                   uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * STILE;
#pragma unroll
    for (int i = 0; i < SITEMS; ++i) {
        size_t idx = base + (size_t)i * SB + threadIdx.x;
        const bool valid = idx < n;
        wave_hist_add(h, valid ? digit_of(keys[idx], bit) : 0u, valid);
    }
    __syncthreads();
    bhist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile (STILE keys).  Keys are ranked inside the tile
// (8 ballots per wave, LDS counters across the 4 waves, rounds in order),
// staged in LDS in digit order, and then written run by run: consecutive
// threads store consecutive addresses of one digit's output run.  The tile's
// count of digit d is the difference of consecutive entries of the scanned
// digit-major histogram.
template <bool HAS_VALS>
__global__ void __launch_bounds__(SB) k_scatter(const Key *kin, const uint32_t *vin, Key *kout, uint32_t *vout,
                                                size_t n, int bit, const uint32_t *boff, uint32_t nblocks) {
    __shared__ uint32_t run_off[256];   // output position of the tile's first key of digit d
    __shared__ uint32_t tile_off[256];  // start of digit d in the staged tile
    __shared__ uint32_t seen[256];      // keys of digit d staged so far
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t scan_tmp[SB / 64 + 1];
    __shared__ Key sk[STILE];
    __shared__ uint32_t sv[HAS_VALS ? STILE : 1];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const size_t hi_idx = (size_t)tid * nblocks + blockIdx.x;
    const uint32_t cur = boff[hi_idx];
    const uint32_t nxt = hi_idx + 1 < (size_t)256 * nblocks ? boff[hi_idx + 1] : (uint32_t)n;
    run_off[tid] = cur;
    uint32_t total;
    tile_off[tid] = block_exclusive_scan(nxt - cur, scan_tmp, total);
    seen[tid] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t base = (size_t)blockIdx.x * STILE;
    for (int r = 0; r < SITEMS; ++r) {
        size_t idx = base + (size_t)r * SB + tid;
        bool valid = idx < n;
        Key k{0ull, 0ull};
        uint32_t v = 0;
        if (valid) {
            k = kin[idx];
            if (HAS_VALS) v = vin[idx];
        }
        uint32_t d = valid ? digit_of(k, bit) : 0u;
#pragma unroll
        for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wid][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = tile_off[d] + seen[d] + rank;
            for (int w = 0; w < wid; ++w) pos += wcnt[w][d];
            sk[pos] = k;
            if (HAS_VALS) sv[pos] = v;
        }
        __syncthreads();
        seen[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    }
    __syncthreads();
    const uint32_t tn = (uint32_t)min((size_t)STILE, n - base);
    for (uint32_t i = tid; i < tn; i += SB) {
        const Key k = sk[i];
        const uint32_t d = digit_of(k, bit);
        const uint32_t pos = run_off[d] + (i - tile_off[d]);
        kout[pos] = k;
        if (HAS_VALS) vout[pos] = sv[i];
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
    HIP_OK(hipMemcpyAsync(&h, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return h;
}

namespace {
void run_passes(Key *keys, uint32_t *vals, size_t n, const std::vector<int> &passes, hipStream_t s);
}  // namespace

void radix_sort_keys(Key *keys, uint32_t *vals, size_t n, hipStream_t s, int begin_bit, int end_bit, bool probe) {
    if (n < 2) return;
    if (n > 0xffffffffull) {
        dgx_set_error("radix_sort_keys: %zu keys exceed the 32-bit index range", n);
        throw DgxError();
    }
    if (!vals && n <= (size_t)SMALL_N) {
        k_sort_small<<<1, 1024, 0, s>>>(keys, (uint32_t)n);
        HIP_OK(hipGetLastError());
        return;
    }
    std::vector<int> passes;
    if (probe) {
        // skip digit positions that are constant over all keys (one host read)
        DevBuf<uint32_t> ghist(16 * 256, s);
        HIP_OK(hipMemsetAsync(ghist.p, 0, 16 * 256 * sizeof(uint32_t), s));
        unsigned gb = (unsigned)std::min<size_t>(grid_for(n, SB), 2048);
        k_global_hist<<<gb, SB, 0, s>>>(keys, n, ghist.p);
        HIP_OK(hipGetLastError());
        std::vector<uint32_t> h(16 * 256);
        HIP_OK(hipMemcpyAsync(h.data(), ghist.p, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        for (int p = 0; p < 16; ++p) {
            int bit = 8 * p;
            if (bit + 8 <= begin_bit || bit >= end_bit) continue;
            bool trivial = false;
            for (int d = 0; d < 256; ++d)
                if (h[p * 256 + d] == n) trivial = true;
            if (!trivial) passes.push_back(bit);
        }
    } else {
        for (int p = 0; p < 16; ++p)
            if (!(8 * p + 8 <= begin_bit || 8 * p >= end_bit)) passes.push_back(8 * p);
    }
    run_passes(keys, vals, n, passes, s);
}

void radix_sort_keys_mask(Key *keys, uint32_t *vals, size_t n, hipStream_t s, uint64_t vary_hi, uint64_t vary_lo) {
    if (n < 2) return;
    if (n > 0xffffffffull) {
        dgx_set_error("radix_sort_keys: %zu keys exceed the 32-bit index range", n);
        throw DgxError();
    }
    if (!vals && n <= (size_t)SMALL_N) {
        k_sort_small<<<1, 1024, 0, s>>>(keys, (uint32_t)n);
        HIP_OK(hipGetLastError());
        return;
    }
    std::vector<int> passes;
    for (int p = 0; p < 16; ++p) {
        const uint64_t w = p < 8 ? vary_lo : vary_hi;
        if ((w >> (8 * (p & 7))) & 255u) passes.push_back(8 * p);
    }
    run_passes(keys, vals, n, passes, s);
}

namespace {

void run_passes(Key *keys, uint32_t *vals, size_t n, const std::vector<int> &passes, hipStream_t s) {
    if (passes.empty()) return;
    size_t nblocks = (n + STILE - 1) / STILE;
    DevBuf<Key> ktmp(n, s);
    DevBuf<uint32_t> vtmp(vals ? n : 0, s);
    DevBuf<uint32_t> bhist(256 * nblocks, s);
    Key *kin = keys, *kout = ktmp.p;
    uint32_t *vin = vals, *vout = vtmp.p;
    for (int bit : passes) {
        k_block_hist<<<(unsigned)nblocks, SB, 0, s>>>(kin, n, bit, bhist.p, (uint32_t)nblocks);
        scan_exclusive_u32_async(bhist.p, bhist.p, 256 * nblocks, nullptr, s);
        if (vals)
            k_scatter<true><<<(unsigned)nblocks, SB, 0, s>>>(kin, vin, kout, vout, n, bit, bhist.p, (uint32_t)nblocks);
        else
            k_scatter<false><<<(unsigned)nblocks, SB, 0, s>>>(kin, nullptr, kout, nullptr, n, bit, bhist.p,
                                                             (uint32_t)nblocks);
        HIP_OK(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        HIP_OK(hipMemcpyAsync(keys, kin, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
        if (vals) HIP_OK(hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace
