// dgx_octree.hip -- points -> linear octree (ot::points2Octree) and 2:1 balance
// (ot::balanceOctree) on gfx950.
//
// points2Octree (src/construct.cpp:911-1077):
//   1. quantise + SFC-encode every point (k_quantise)            [construct.cpp:929-945]
//   2. radix-sort the 90-bit keys                                [omp_par::merge_sort, :955]
//   3. "textbook" leaves: an octant is split iff it holds more than maxNumPts
//      points.  O is split iff it contains some pair (p_i, p_{i+m}), so the
//      leaf level of point j is 1 + max over windows i in [j-m, j] of the
//      common-prefix level of (p_i, p_{i+m}) (k_p2o_lev, k_p2o_leaf).
//   4. empty leaves fill the gaps between consecutive non-empty leaves as the
//      coarsest octants of the gap, enumerated in digit space (k_gap_*).
//   5. p2oLocal (:976-1077) advances a point cursor through windows of
//      maxNumPts points; it equals the textbook tree except after level-maxDepth
//      cells whose points do not fit in one window (duplicate clusters).  Those
//      cells are detected (k_p2o_trigger) and the reference state machine is
//      replayed exactly from there (k_p2o_replay) until it is back in step.
// balanceOctree (src/balance.cpp:55-64 -> comboRipple; TreeNode::balanceSubtree
// TreeNode.cpp:884-1343): level by level from the finest, the parents of all
// seeds are made unique (their 8 children are in the balanced tree) and their
// 26 neighbours seed the next coarser level.  Leaves are the children of split
// octants that are not split.  Per-level work uses Morton keys; the final leaf
// set is re-keyed on the output curve and sorted once.
#include "dgx_octree.h"

#include <algorithm>
#include <vector>

namespace {

#define LOAD_TABLES(tab)                                                 \
    __shared__ uint8_t s_rc[192], s_mc[192];                             \
    for (int i_ = threadIdx.x; i_ < 192; i_ += blockDim.x) {             \
        s_rc[i_] = (tab).rc[i_];                                          \
        s_mc[i_] = (tab).mc[i_];                                          \
    }                                                                    \
    __syncthreads();
// LOAD_TABLES plus the two-level encode table s_rc2 (morton_to_curve_rot).
#define LOAD_TABLES2(tab)                                                \
    __shared__ uint8_t s_rc[192], s_mc[192];                             \
    __shared__ uint16_t s_rc2[24 * 64];                                  \
    for (int i_ = threadIdx.x; i_ < 192; i_ += blockDim.x) {             \
        s_rc[i_] = (tab).rc[i_];                                          \
        s_mc[i_] = (tab).mc[i_];                                          \
    }                                                                    \
    for (int i_ = threadIdx.x; i_ < 24 * 64; i_ += blockDim.x)           \
        s_rc2[i_] = rc2_entry((tab).rc, (uint32_t)i_);                    \
    __syncthreads();

constexpr int B = 256;

// Level-maxDepth keys of the points; *bad is set when a point lies outside
// the domain.
__global__ void __launch_bounds__(B) k_quantise(const double *xyz, size_t n, uint32_t md, double gx, double gy,
                                                double gz, SfcPacked tab, Key *keys, uint32_t *bad) {
    LOAD_TABLES(tab)
    const size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double scale = (double)(1u << md);
    // (unsigned)(x * 2^maxDepth / gLen), construct.cpp:933-941
    const double fx = xyz[3 * i] * scale / gx, fy = xyz[3 * i + 1] * scale / gy, fz = xyz[3 * i + 2] * scale / gz;
    uint32_t px = 0, py = 0, pz = 0;
    const bool inside = fx >= 0.0 && fx < scale && fy >= 0.0 && fy < scale && fz >= 0.0 && fz < scale;
    if (inside) {
        px = (uint32_t)fx;
        py = (uint32_t)fy;
        pz = (uint32_t)fz;
    } else {
        atomicOr(bad, 1u);
    }
    keys[i] = dev_encode(s_rc, px, py, pz, md, md);
}

__global__ void k_p2o_lev(const Key *pk, size_t n, uint32_t m, uint32_t md, uint8_t *lev) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i + m >= n) return;
    uint32_t c = key_common_levels(pk[i], pk[i + m]);
    lev[i] = (uint8_t)min(c, md - 1);
}

__global__ void k_p2o_leaf(const Key *pk, size_t n, uint32_t m, const uint8_t *lev, Key *leafkey) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    size_t lo = j >= m ? j - m : 0, hi = min(j, n - m - 1);
    uint32_t L = 0;
    for (size_t i = lo; i <= hi; ++i) L = max(L, (uint32_t)lev[i]);
    leafkey[j] = key_truncate(pk[j], L + 1);
}

// Keys sorted on their high words alone: each run of equal high words is put
// in order of bits 8-63 of the low word (stably), which completes a stable
// sort on bits 8-127.  The wave that holds a run's first key sorts the run:
// lane l ranks keys l, l+64, ... against the whole run, then stores them at
// their ranks.  Only low words move, and only that wave reads them, so the
// high words other waves compare stay put.  A run longer than kHiRunCap is
// left alone and sets *longrun.
constexpr uint32_t kHiRunCap = 256, kHiRunQ = kHiRunCap / 64;
__global__ void __launch_bounds__(B) k_sort_hi_runs(Key *k, size_t n, uint32_t *longrun) {
    const size_t j = (size_t)blockIdx.x * B + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    // one load per key; the neighbours' high words come from the adjacent
    // lanes, and from memory at the ends of the wave
    const uint64_t h = j < n ? k[j].hi : 0ull;
    uint64_t hp = __shfl_up(h, 1), hn = __shfl_down(h, 1);
    if (lane == 0 && j > 0 && j <= n) hp = k[j - 1].hi;
    if (lane == 63 && j + 1 < n) hn = k[j + 1].hi;
    const bool start = j + 1 < n && hn == h && !(j > 0 && hp == h);
    uint64_t starts = __ballot(start);
    const size_t wbase = j - lane;
    while (starts) {
        const uint32_t sl = (uint32_t)__ffsll((unsigned long long)starts) - 1;
        starts &= starts - 1;
        const size_t j0 = wbase + sl;
        const uint64_t h0 = __shfl(h, (int)sl);
        // run end: the first key past j0 with another high word
        size_t e = n;
        for (size_t c = j0 + 1; c < n; c += 64) {
            const size_t idx = c + lane;
            const uint64_t other = __ballot(idx < n && k[idx].hi != h0);
            if (other) {
                e = c + (size_t)__ffsll((unsigned long long)other) - 1;
                break;
            }
            if (c + 64 - j0 > kHiRunCap) break;
        }
        const size_t len = e - j0;
        if (len > kHiRunCap) {
            if (lane == 0) atomicOr(longrun, 1u);
            continue;
        }
        uint64_t v[kHiRunQ];
        uint32_t r[kHiRunQ];
#pragma unroll
        for (uint32_t q = 0; q < kHiRunQ; ++q) {
            const uint32_t t = lane + 64 * q;
            v[q] = t < len ? k[j0 + t].lo : 0ull;
            r[q] = 0;
        }
        for (uint32_t t = 0; t < len; ++t) {
            const uint64_t w = k[j0 + t].lo >> 8;
#pragma unroll
            for (uint32_t q = 0; q < kHiRunQ; ++q) {
                const uint64_t x = v[q] >> 8;
                r[q] += (w < x || (w == x && t < lane + 64 * q)) ? 1u : 0u;
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kHiRunQ; ++q)
            if (lane + 64 * q < len) k[j0 + r[q]].lo = v[q];
    }
}

__global__ void k_head(const Key *k, size_t n, uint32_t *head) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    head[j] = (j == 0 || !key_eq(k[j], k[j - 1])) ? 1u : 0u;
}

__global__ void k_compact_idx(const uint32_t *flag, const uint32_t *pos, size_t n, uint32_t *out) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (flag[j]) out[pos[j]] = (uint32_t)j;
}

// Number of octants in the coarsest completion strictly between A and B (A < B disjoint).
__device__ inline uint32_t gap_between(const Key &A, const Key &Bk) {
    uint32_t k = key_common_levels(A, Bk), la = key_level(A), lb = key_level(Bk);
    uint32_t c = key_digit(Bk, k + 1) - key_digit(A, k + 1) - 1;
    for (uint32_t t = k + 2; t <= la; ++t) c += 7 - key_digit(A, t);
    for (uint32_t t = k + 2; t <= lb; ++t) c += key_digit(Bk, t);
    return c;
}
__device__ inline uint32_t gap_before(const Key &Bk) {
    uint32_t c = 0, lb = key_level(Bk);
    for (uint32_t t = 1; t <= lb; ++t) c += key_digit(Bk, t);
    return c;
}
__device__ inline uint32_t gap_after(const Key &A) {
    uint32_t c = 0, la = key_level(A);
    for (uint32_t t = 1; t <= la; ++t) c += 7 - key_digit(A, t);
    return c;
}

// q in [0, Q]: count of octants emitted by slot q (the gap before leaf q plus the leaf; slot Q: the tail gap).
// Consecutive non-empty leaves must be disjoint and increasing; otherwise the
// slot counts 0 and err is raised (the host then refuses the result).
// cnt[q] for q <= cap: octants emitted for leaf q (0 beyond the device count
// *dQ of distinct leaves, so the scan may run over the capacity).
__global__ void k_gap_count(const Key *leafkey, const uint32_t *NL, const uint32_t *dQ, size_t cap, uint32_t *cnt,
                            uint32_t *err) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > cap) return;
    const size_t Q = *dQ;
    if (q > Q) {
        cnt[q] = 0;
        return;
    }
    if (q == Q) {
        cnt[q] = gap_after(leafkey[NL[Q - 1]]);
    } else {
        Key b = leafkey[NL[q]];
        if (q == 0) {
            cnt[q] = 1 + gap_before(b);
        } else {
            Key a = leafkey[NL[q - 1]];
            uint32_t k = key_common_levels(a, b);
            if (k >= min(key_level(a), key_level(b)) || key_digit(b, k + 1) <= key_digit(a, k + 1)) {
                atomicOr(err, 1u);
                cnt[q] = 0;
                return;
            }
            cnt[q] = 1 + gap_between(a, b);
        }
    }
}

__device__ inline Key with_digit(const Key &base, uint32_t t, uint32_t r) {
    Key k = key_truncate(base, t);
    key_set_digit(k, t, r);
    return k;
}

__global__ void k_gap_emit(const Key *leafkey, const uint32_t *NL, const uint32_t *dQ, size_t n, const uint32_t *off,
                           Key *out, uint32_t *ptb, uint32_t *ptc) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t Q = *dQ;
    if (q > Q) return;
    uint32_t o = off[q];
    if (q == Q) {
        Key A = leafkey[NL[Q - 1]];
        uint32_t la = key_level(A);
        for (uint32_t t = la; t >= 1; --t)
            for (uint32_t r = key_digit(A, t) + 1; r < 8; ++r) {
                out[o] = with_digit(A, t, r);
                ptb[o] = (uint32_t)n;
                ptc[o] = 0;
                ++o;
            }
        return;
    }
    uint32_t b0 = NL[q], b1 = (q + 1 < Q) ? NL[q + 1] : (uint32_t)n;
    Key Bk = leafkey[b0];
    uint32_t lb = key_level(Bk), t0 = 1;
    if (q > 0) {
        Key A = leafkey[NL[q - 1]];
        uint32_t k = key_common_levels(A, Bk), la = key_level(A);
        for (uint32_t t = la; t >= k + 2; --t)
            for (uint32_t r = key_digit(A, t) + 1; r < 8; ++r) {
                out[o] = with_digit(A, t, r); ptb[o] = b0; ptc[o] = 0; ++o;
            }
        for (uint32_t r = key_digit(A, k + 1) + 1; r < key_digit(Bk, k + 1); ++r) {
            out[o] = with_digit(A, k + 1, r); ptb[o] = b0; ptc[o] = 0; ++o;
        }
        t0 = k + 2;
    }
    for (uint32_t t = t0; t <= lb; ++t)
        for (uint32_t r = 0; r < key_digit(Bk, t); ++r) {
            out[o] = with_digit(Bk, t, r); ptb[o] = b0; ptc[o] = 0; ++o;
        }
    out[o] = Bk;
    ptb[o] = b0;
    ptc[o] = b1 - b0;
}

// Arrival cursor of the reference state machine at the level-maxDepth leaf k,
// assuming the leaves before it were emitted in step (see file comment).
__device__ inline uint64_t arrival_cursor(const Key *T, const uint32_t *tb, const uint32_t *tc, size_t k,
                                          uint32_t md) {
    uint64_t s = tb[k];
    if (k > 0 && key_level(T[k - 1]) == md) {
        Key nx;
        if (key_next(T[k - 1], nx) && key_eq(nx, T[k])) s += tc[k];
    }
    return s;
}

__global__ void k_p2o_trigger(const Key *T, const uint32_t *tb, const uint32_t *tc, size_t nT, size_t n, uint32_t m,
                              uint32_t md, uint32_t *flag) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nT) return;
    uint32_t f = 0;
    if (key_level(T[k]) == md) {
        uint64_t s = arrival_cursor(T, tb, tc, k, md);
        if (s + m < n) {
            uint64_t E = (uint64_t)tb[k] + tc[k];
            Key nx;
            if (key_next(T[k], nx) && key_level(nx) == md && k + 1 < nT) E += tc[k + 1];
            f = (E > s + m) ? 1u : 0u;
        }
    }
    flag[k] = f;
}

__device__ inline size_t lower_bound_lt(const Key *a, size_t lo, size_t hi, const Key &v) {  // first a[i] >= v
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_lt(a[mid], v)) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ inline size_t lower_bound_le(const Key *a, size_t lo, size_t hi, const Key &v) {  // first a[i] > v
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_le(a[mid], v)) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Exact replay of p2oLocal (construct.cpp:1000-1061) from the state after
// the trigger leaf k, on one thread, until the machine is back in step with the
// textbook leaves at an octant coarser than maxDepth.
__global__ void k_p2o_replay(const Key *P, size_t n, uint32_t m, uint32_t md, const Key *T, size_t nT,
                             const uint32_t *tb, const uint32_t *tc, size_t ktrig, Key *out, uint32_t cap,
                             ReplayResult *res) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ReplayResult r;
    r.nout = 0;
    r.end = 0;
    r.err = 0;
    r.resync = Key{0, 0};
    r.next_t = (uint32_t)nT;
    uint64_t s = arrival_cursor(T, tb, tc, ktrig, md);
    Key curr;
    if (!key_next(T[ktrig], curr)) {
        r.end = 1;
        *res = r;
        return;
    }
    uint64_t curr_pt = s + m, next_pt = curr_pt + m;
    Key next;
    bool has_next = key_next(curr, next);
    bool synced = false;
    while (next_pt < n) {
        if (key_level(curr) < md && curr_pt == lower_bound_lt(P, 0, n, curr)) {
            size_t j = lower_bound_lt(T, 0, nT, curr);
            if (j < nT && (key_eq(T[j], curr) || key_is_ancestor(curr, T[j]))) {
                synced = true;
                break;
            }
        }
        while (key_is_ancestor(curr, P[next_pt - 1]) && key_is_ancestor(curr, P[next_pt]) && key_level(curr) < md) {
            curr = key_first_child(curr);
            has_next = key_next(curr, next);
        }
        if (key_level(curr) == md - 1) next_pt = n - 1;  // construct.cpp:1022-1032
        if (!has_next) {  // the reference would compare points with the past-the-end sentinel
            r.err = 2;
            *res = r;
            return;
        }
        uint64_t found = lower_bound_le(P, curr_pt, next_pt, next) - curr_pt;
        next_pt = curr_pt + found;
        if (r.nout >= cap) { r.err = 1; *res = r; return; }
        out[r.nout++] = curr;
        curr = next;
        has_next = key_next(curr, next);
        if (next_pt > curr_pt) curr_pt = next_pt;
        next_pt = curr_pt + m;
    }
    if (synced) {
        r.resync = curr;
        r.next_t = (uint32_t)lower_bound_lt(T, ktrig + 1, nT, curr);  // first textbook leaf in resync
    } else {  // tail loop (:1051-1061): coarsest completion to the end of the domain
        while (true) {
            if (r.nout >= cap) { r.err = 1; *res = r; return; }
            out[r.nout++] = curr;
            Key nx;
            if (!key_next(curr, nx)) break;
            curr = nx;
        }
        r.end = 1;
    }
    *res = r;
}

// ------------------------------------------------------------ balance ----
// Histogram of the input levels, 8 keys per thread: few blocks, so few global
// atomics on the 32 counters.
constexpr int LH_ITEMS = 8;
__global__ void __launch_bounds__(B) k_level_hist(const Key *k, size_t n, uint32_t *hist) {
    __shared__ uint32_t h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * B * LH_ITEMS + threadIdx.x;
#pragma unroll
    for (int r = 0; r < LH_ITEMS; ++r) {
        const size_t i = base + (size_t)r * B;
        wave_hist_add(h, i < n ? key_level(k[i]) : 0u, i < n);
    }
    __syncthreads();
    if (threadIdx.x < 32 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Parents (as Morton keys) of the input octants, bucketed by the input level.
// Each block (8 keys per thread) reserves its share of every level bucket
// with one atomic per level; the order inside a bucket is free.
__global__ void __launch_bounds__(B) k_bucket_parents(const Key *k, size_t n, uint32_t md, SfcPacked tab,
                                                      uint32_t *cursor /*32, pre-offset*/, Key *out) {
    LOAD_TABLES(tab)
    __shared__ uint32_t cnt[32], base[32];
    if (threadIdx.x < 32) cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * B * LH_ITEMS + threadIdx.x;
    uint32_t loc[LH_ITEMS], lv[LH_ITEMS];
    Key par[LH_ITEMS];
#pragma unroll
    for (int r = 0; r < LH_ITEMS; ++r) {
        const size_t i = i0 + (size_t)r * B;
        uint32_t x = 0, y = 0, z = 0, l = 0;
        if (i < n) dev_decode(s_mc, k[i], md, x, y, z, l);
        loc[r] = wave_bucket_slot(cnt, l, i < n);
        lv[r] = l;
        par[r] = key_parent(morton_encode(x, y, z, l, md));
    }
    __syncthreads();
    if (threadIdx.x < 32) base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]) : 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < LH_ITEMS; ++r)
        if (i0 + (size_t)r * B < n) out[base[lv[r]] + loc[r]] = par[r];
}

// Parents of the 26 neighbours of every split octant: per axis the neighbours'
// parents take at most two values, so at most 8 distinct parents per octant.
// They go after the next level's nseed input parents, which the same launch
// copies to out[0..nseed).
// Neighbour parents of the split octants Q (level lvl): for each octant, per
// axis its own parent coordinate and the neighbouring one on the side of the
// octant inside its parent; the up-to-8 combinations inside the domain are
// parents at level lvl-1 of the octant's 26 neighbours.
__device__ __forceinline__ uint32_t nbr_parent_axes(const Key &q, uint32_t md, uint32_t lvl, uint32_t c0[3],
                                                    uint32_t c1[3]) {
    uint32_t p[3], l;
    morton_decode(q, md, p[0], p[1], p[2], l);
    const uint32_t sh = md - lvl + 1;  // log2 of the parent length
    const uint64_t plen = 1ull << sh, dom = 1ull << md;
    uint32_t ok[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint64_t lo = (uint64_t)(p[a] >> sh) << sh;
        const bool upper = (p[a] & (uint32_t)(plen - 1)) != 0;  // the octant is in the upper half of its parent
        c0[a] = (uint32_t)lo;
        const uint64_t nb = upper ? lo + plen : lo - plen;  // wraps below 0: rejected by the domain test
        c1[a] = (uint32_t)nb;
        ok[a] = 1u | ((upper ? nb < dom : lo >= plen) ? 2u : 0u);
    }
    uint32_t valid = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t)
        if (((ok[0] >> (t & 1)) & 1u) && ((ok[1] >> ((t >> 1) & 1)) & 1u) && ((ok[2] >> ((t >> 2) & 1)) & 1u))
            valid |= 1u << t;
    return valid;
}

// The next level's seeds: the input parents (copied), then the neighbour
// parents of the split octants, NB_ITEMS octants per thread and one atomic
// per workgroup to reserve the output.
constexpr int NB_ITEMS = 4;
__global__ void __launch_bounds__(B) k_bal_nbr_parents(const Key *Q, const uint32_t *d_nq, uint32_t lvl, uint32_t md,
                                                       Key *out, uint32_t *counter, const Key *seed, size_t nseed) {
    __shared__ uint32_t red[B / 64 + 1];
    const size_t i0 = (size_t)blockIdx.x * B * NB_ITEMS + threadIdx.x;
    const uint32_t nq = *d_nq;
    uint32_t c0[NB_ITEMS][3], c1[NB_ITEMS][3], valid[NB_ITEMS];
    uint32_t cnt = 0;
#pragma unroll
    for (int r = 0; r < NB_ITEMS; ++r) {
        const size_t i = i0 + (size_t)r * B;
        if (i < nseed) out[i] = seed[i];
        valid[r] = 0;
        if (i < nq) valid[r] = nbr_parent_axes(Q[i], md, lvl, c0[r], c1[r]);
        cnt += (uint32_t)__popc(valid[r]);
    }
    out += nseed;
    uint32_t base = block_reserve(counter, cnt, red);
#pragma unroll
    for (int r = 0; r < NB_ITEMS; ++r)
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if ((valid[r] >> t) & 1u)
                out[base++] = morton_encode((t & 1) ? c1[r][0] : c0[r][0], (t & 2) ? c1[r][1] : c0[r][1],
                                            (t & 4) ? c1[r][2] : c0[r][2], lvl - 1, md);
}

// For every split octant at the finer level, mark its child bit in the parent (Qp sorted unique).
__global__ void k_mark_split_children(const Key *Qc, size_t nc, const Key *Qp, const uint32_t *d_np,
                                      uint32_t *cmask) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const size_t np = *d_np;
    Key par = key_parent(Qc[i]);
    size_t j = lower_bound_lt(Qp, 0, np, par);
    if (j >= np || !key_eq(Qp[j], par)) return;  // cannot happen: split octants are closed under parents
    uint32_t d = key_digit(Qc[i], key_level(Qc[i]));
    atomicOr(&cmask[j], 1u << d);
}

// cnt[i] for i < cap: leaves under split octant i (0 beyond the device count).
// Single-pass forms of the balance level's two compactions (flag, scan,
// write at offset): tiles of FS_TILE entries, FS_ITEMS consecutive entries
// per thread, tile ids from ctr[epoch], offsets by decoupled look-back.
constexpr uint32_t FS_B = 256, FS_ITEMS = 16, FS_TILE = FS_B * FS_ITEMS;

// Q = the distinct keys of the sorted W[0..n) in order, *d_nq their count;
// also clears cmask[0..*d_nq) and *zcount for the kernels that follow.
__global__ void __launch_bounds__(FS_B) k_unique_compact(const Key *W, size_t n, Key *Q, uint32_t *cmask,
                                                         uint32_t *d_nq, uint32_t *zcount, uint64_t *status,
                                                         uint32_t *ctr, uint32_t epoch, uint32_t *fault) {
    __shared__ uint32_t s_tile, s_excl, tmp[FS_B / 64 + 1];
    if (threadIdx.x == 0) s_tile = atomicAdd(&ctr[epoch], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t i0 = (size_t)tile * FS_TILE + (size_t)threadIdx.x * FS_ITEMS;
    Key kv[FS_ITEMS];
    Key prev = (i0 > 0 && i0 - 1 < n) ? W[i0 - 1] : Key{~0ull, ~0ull};
    uint32_t fl = 0, cnt = 0;
#pragma unroll
    for (uint32_t e = 0; e < FS_ITEMS; ++e) {
        const size_t i = i0 + e;
        if (i < n) {
            kv[e] = W[i];
            if (i == 0 || !key_eq(kv[e], prev)) {
                fl |= 1u << e;
                ++cnt;
            }
            prev = kv[e];
        }
    }
    uint32_t total;
    const uint32_t tex = block_excl_scan_u32(cnt, tmp, total);
    if (threadIdx.x == 0) s_excl = lookback_prefix(status, tile, total, epoch, fault);
    __syncthreads();
    uint32_t pos = s_excl + tex;
#pragma unroll
    for (uint32_t e = 0; e < FS_ITEMS; ++e)
        if ((fl >> e) & 1u) {
            Q[pos] = kv[e];
            cmask[pos] = 0;
            ++pos;
        }
    if (threadIdx.x == 0) {
        if (tile + 1 == gridDim.x) *d_nq = s_excl + total;
        if (tile == 0) *zcount = 0;
    }
}

// The leaves of one level: the children of the split octants Q[0..*d_nq) that
// are not split themselves (cmask), in Q order and child order, written to
// out from offset 0; *d_nlv their count.  Tiles of 2048 octants on 1024
// threads, 2 octants each: emitting an octant's leaves is a chain of table
// reads, so fewer octants per thread keep the chains short.
constexpr uint32_t LE_B = 1024, LE_ITEMS = 2, LE_TILE = LE_B * LE_ITEMS;
__global__ void __launch_bounds__(LE_B) k_leaf_emit_scan(const Key *Q, const uint32_t *d_nq, const uint32_t *cmask,
                                                         uint32_t md, SfcPacked tab, Key *out, uint32_t *d_nlv,
                                                         uint64_t *status, uint32_t *ctr, uint32_t epoch,
                                                         uint32_t *fault) {
    LOAD_TABLES2(tab)
    __shared__ uint32_t s_tile, s_excl, tmp[LE_B / 64 + 1];
    if (threadIdx.x == 0) s_tile = atomicAdd(&ctr[epoch], 1u);
    __syncthreads();
    const uint32_t tile = s_tile, nq = *d_nq;
    const size_t i0 = (size_t)tile * LE_TILE + (size_t)threadIdx.x * LE_ITEMS;
    uint32_t msk[LE_ITEMS], cnt = 0;
#pragma unroll
    for (uint32_t e = 0; e < LE_ITEMS; ++e) {
        const size_t i = i0 + e;
        msk[e] = i < nq ? cmask[i] : 0xffu;
        cnt += 8u - (uint32_t)__popc(msk[e] & 255u);
    }
    uint32_t total;
    const uint32_t tex = block_excl_scan_u32(cnt, tmp, total);
    if (threadIdx.x == 0) s_excl = lookback_prefix(status, tile, total, epoch, fault);
    __syncthreads();
    uint32_t pos = s_excl + tex;
    for (uint32_t e = 0; e < LE_ITEMS; ++e) {
        const size_t i = i0 + e;
        if (i >= nq) break;
        uint32_t rot;
        const Key pk = morton_to_curve_rot(s_rc2, s_rc, Q[i], rot);
        for (uint32_t c = 0; c < 8; ++c)
            if (!((msk[e] >> c) & 1u)) out[pos++] = child_key(s_rc, pk, rot, c);
    }
    if (threadIdx.x == 0 && tile + 1 == gridDim.x) *d_nlv = s_excl + total;
}

__global__ void __launch_bounds__(B) k_export_tn(const Key *k, size_t n, uint32_t md, SfcPacked tab, uint32_t *tn) {
    LOAD_TABLES(tab)
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z, l;
    dev_decode(s_mc, k[i], md, x, y, z, l);
    uint32_t *o = tn + 7 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = l; o[4] = 1; o[5] = 3; o[6] = md;
}

__global__ void __launch_bounds__(B) k_import_tn(const uint32_t *tn, size_t n, uint32_t md, SfcPacked tab, Key *k,
                                                 uint32_t *bad) {
    LOAD_TABLES(tab)
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const uint32_t *t = tn + 7 * i;
    uint32_t l = t[3] & 31u;
    uint32_t x = t[0], y = t[1], z = t[2];
    if (l > md || (l < md && ((x | y | z) & ((1u << (md - l)) - 1u))) || (x >> md) || (y >> md) || (z >> md))
        atomicOr(bad, 1u);
    k[i] = dev_encode(s_rc, x, y, z, l, md);
}

// ---- consecutive small balance levels in one workgroup -------------------
// The level step of balance_build for levels whose seeds fit kFuseCap keys,
// run by one workgroup for as many consecutive levels as stay that small:
// sort + unique the seeds (bitonic in LDS), mark the children that are split
// themselves, emit the leaves, and collect the next level's seeds (the input
// parents of that level, then the neighbour parents).  Leaves and neighbour
// parents may come out in any order: seeds are sorted per level and the
// leaves once at the end, exactly as in the multi-kernel step.
// 4096 seeds (4 per thread) ran the same build in the same time: the levels
// it adds cost the single workgroup what their kernels and host reads cost.
constexpr uint32_t kFuseCap = 2048, kFuseB = 1024, kFuseF = kFuseCap / kFuseB;  // seeds per thread
struct FuseState {
    uint32_t level;      // next level to process (0: all done)
    uint32_t nw;         // its seeds: in wbuf[wsel], or the input parents of that level when from_inpar
    uint32_t wsel, from_inpar;
    uint32_t nprev;      // split octants of the last processed level, in qbuf[qsel]
    uint32_t qsel;
    uint32_t used;       // leaves emitted so far
    uint32_t levels, seeds;
};

__device__ __forceinline__ uint32_t fuse_block_scan(uint32_t v, uint32_t *tmp, uint32_t &total) {
    // exclusive scan over the kFuseB threads; tmp holds kFuseB/64 + 1 entries
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)lane >= o) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < kFuseB / 64; ++w) {
            const uint32_t t = tmp[w];
            tmp[w] = acc;
            acc += t;
        }
        tmp[kFuseB / 64] = acc;
    }
    __syncthreads();
    const uint32_t r = x - v + tmp[wid];
    total = tmp[kFuseB / 64];
    __syncthreads();
    return r;
}

// Orders this wave's LDS accesses before and after the call (LDS operations
// of one wave complete in order); no other wave is waited for.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ bool fuse_less(const uint64_t &a, const uint64_t &b) { return a < b; }
__device__ __forceinline__ bool fuse_less(const Key &a, const Key &b) { return key_lt(a, b); }

// Bitonic sort of a[0..P) (P a power of two <= kFuseCap) by the kFuseB
// threads.  In every stage thread t performs compare-exchanges c = t + q kFuseB:
// elements i = c with a zero bit inserted at log2(j), and i + j.  For partner
// distances j <= 64 the exchanges of wave w touch only elements
// [128 w + 2 kFuseB q, +128), which no other wave touches, so consecutive such
// stages are separated by a wave-level fence only.
template <typename T>
__device__ void fuse_bitonic(T *a, uint32_t P) {
    const uint32_t t = threadIdx.x, ncmp = P >> 1;
    bool prev_local = false;
    for (uint32_t kk = 2; kk <= P; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            const bool local = j <= 64;
            if (local && prev_local)
                wave_lds_sync();
            else
                __syncthreads();
            const uint32_t lj = __ffs(j) - 1;
            for (uint32_t c = t; c < ncmp; c += kFuseB) {
                const uint32_t i = ((c >> lj) << (lj + 1)) | (c & (j - 1)), o = i + j;
                const bool up = (i & kk) == 0;
                const T x = a[i], y = a[o];
                if (up ? fuse_less(y, x) : fuse_less(x, y)) {
                    a[i] = y;
                    a[o] = x;
                }
            }
            prev_local = local;
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kFuseB) k_bal_fused(uint32_t l0, const Key *win, uint32_t nw0, const Key *prevq,
                                                     uint32_t nprev0, const Key *inpar, const uint32_t *loff,
                                                     uint32_t md, SfcPacked tab, Key *leaves, uint32_t used0,
                                                     Key *wbuf0, Key *wbuf1, Key *qbuf0, Key *qbuf1, FuseState *st) {
    __shared__ uint32_t s_loff[33];  // input parents of level l: inpar[s_loff[l - 1] .. s_loff[l])
    for (uint32_t i = threadIdx.x; i < 33; i += kFuseB) s_loff[i] = loff[i];
    LOAD_TABLES2(tab)
    __shared__ Key sk[kFuseCap];
    __shared__ uint64_t sh[kFuseCap];  // the high words, when the low word is the level alone
    __shared__ uint32_t cm[kFuseCap];
    __shared__ uint32_t tmp[kFuseB / 64 + 1];
    // while the run stays in this kernel, the next level's seeds (wl) and this
    // level's split octants (pq) stay in LDS; global memory gets them only
    // when the host takes over
    __shared__ Key wl[kFuseCap];
    __shared__ Key pq[kFuseCap];
    const uint32_t t = threadIdx.x;
    uint32_t l = l0, nw = nw0, nprev = nprev0, used = used0, levels = 0, seeds = 0, k = 0;
    const Key *W = win, *PQ = prevq;
    bool from_inpar = false, w_lds = false, pq_lds = false;
    for (;;) {
        // The next level's input parents are loaded now (one per thread; the
        // rest when they are stored), so the load overlaps this level's work.
        const uint32_t nin_next = l >= 2 ? s_loff[l] - s_loff[l - 1] : 0u;
        Key pf{0ull, 0ull};
        if (t < nin_next) pf = inpar[s_loff[l - 1] + t];
        // seeds of level l (1 <= nw <= kFuseCap): sort, then unique in place -> Q = sk[0..nq)
        uint32_t P = 1;
        while (P < nw) P <<= 1;
        if (P <= 64) {
            // a few seeds (the staircase below a cluster of identical points):
            // one wave sorts them in registers by a bitonic network of shuffles
            if (t < 64) {
                Key w = t < nw ? (w_lds ? wl[t] : W[t]) : Key{~0ull, ~0ull};
                for (uint32_t kk = 2; kk <= P; kk <<= 1)
                    for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                        Key o;
                        o.hi = __shfl_xor(w.hi, (int)j, 64);
                        o.lo = __shfl_xor(w.lo, (int)j, 64);
                        const bool up = (t & kk) == 0, lower = (t & j) == 0;
                        if (lower == up ? key_lt(o, w) : key_lt(w, o)) w = o;
                    }
                if (t < nw) sk[t] = w;
            }
        } else {
            // seeds are parents at level l-1: at levels <= 21 every digit is in the
            // high word and the low word is the level, so 8-byte values are sorted
            const uint64_t lo_lv = l - 1;
            bool narrow = true;
            for (uint32_t i = t; i < P; i += kFuseB) {
                const Key w = i < nw ? (w_lds ? wl[i] : W[i]) : Key{~0ull, lo_lv};
                sh[i] = w.hi;
                narrow = narrow && w.lo == lo_lv;
            }
            narrow = __syncthreads_and(narrow) != 0;
            if (narrow) {
                fuse_bitonic(sh, P);
                for (uint32_t i = t; i < nw; i += kFuseB) sk[i] = Key{sh[i], lo_lv};
            } else {
                for (uint32_t i = t; i < P; i += kFuseB) sk[i] = i < nw ? (w_lds ? wl[i] : W[i]) : Key{~0ull, ~0ull};
                fuse_bitonic(sk, P);
            }
        }
        __syncthreads();
        // thread t owns entries kFuseF t .. kFuseF t + kFuseF - 1
        const uint32_t ib = kFuseF * t;
        Key kq[kFuseF];
        uint32_t fq[kFuseF], fsum = 0;
#pragma unroll
        for (uint32_t e = 0; e < kFuseF; ++e) {
            const uint32_t i = ib + e;
            kq[e] = sk[i];
            fq[e] = i < nw && (i == 0 || !key_eq(kq[e], e ? kq[e - 1] : sk[i - 1])) ? 1u : 0u;
            fsum += fq[e];
        }
        uint32_t nq;
        uint32_t qpos = fuse_block_scan(fsum, tmp, nq);  // also orders the reads above before the writes
#pragma unroll
        for (uint32_t e = 0; e < kFuseF; ++e)
            if (fq[e]) sk[qpos++] = kq[e];
        __syncthreads();
        Key *Qg = (k & 1) ? qbuf1 : qbuf0;  // this level's split octants, for the host
        for (uint32_t i = t; i < nq; i += kFuseB) cm[i] = 0;
        __syncthreads();
        // children of Q that are split octants one level finer
        for (uint32_t i = t; i < nprev; i += kFuseB) {
            const Key c = pq_lds ? pq[i] : PQ[i], par = key_parent(c);
            const size_t j = lower_bound_lt(sk, 0, nq, par);
            if (j < nq && key_eq(sk[j], par)) atomicOr(&cm[j], 1u << key_digit(c, key_level(c)));
        }
        __syncthreads();
        // leaves of level l: the children of Q that are not split
        uint32_t csum = 0;
#pragma unroll
        for (uint32_t e = 0; e < kFuseF; ++e) csum += ib + e < nq ? 8u - __popc(cm[ib + e] & 255u) : 0u;
        uint32_t nlv;
        uint32_t lpos = used + fuse_block_scan(csum, tmp, nlv);
        for (uint32_t e = 0; e < kFuseF; ++e) {
            const uint32_t i = ib + e;
            if (i >= nq) continue;
            uint32_t rot;
            const Key pk = morton_to_curve_rot(s_rc2, s_rc, sk[i], rot);
            const uint32_t msk = cm[i];
            for (uint32_t c = 0; c < 8; ++c) {
                if (msk & (1u << c)) continue;
                leaves[lpos++] = child_key(s_rc, pk, rot, c);
            }
        }
        used += nlv;
        levels += 1;
        seeds += nw;
        if (l < 2) {
            l = 0;
            break;
        }
        // seeds of level l-1: input parents of that level, then neighbour parents of Q
        const uint32_t nin = nin_next;
        const Key *src = inpar + s_loff[l - 1];
        Key *Wn = (k & 1) ? wbuf1 : wbuf0;
        const uint32_t lvl = l - 1;
        uint32_t pc0[kFuseF][3], pc1[kFuseF][3];
        uint32_t valid[kFuseF], vsum = 0;
#pragma unroll
        for (uint32_t e = 0; e < kFuseF; ++e) {
            const uint32_t i = ib + e;
            valid[e] = i < nq ? nbr_parent_axes(sk[i], md, lvl, pc0[e], pc1[e]) : 0u;
            vsum += (uint32_t)__popc(valid[e]);
        }
        uint32_t nnb;
        uint32_t npos = nin + fuse_block_scan(vsum, tmp, nnb);
        const uint32_t nwn = nin + nnb;  // <= largest level input count + 8 kFuseCap = the buffers' capacity
        // level l-1 runs here too when it has seeds and they fit (a level
        // without seeds starts from input parents, below, and needs no Q)
        const bool stay = nwn > 0 && nwn <= kFuseCap;
        Key *wdst = stay ? wl : Wn;
        if (t < nin) wdst[t] = pf;
        for (uint32_t i = t + kFuseB; i < nin; i += kFuseB) wdst[i] = src[i];
        if (stay) {
            for (uint32_t i = t; i < nq; i += kFuseB) pq[i] = sk[i];
        } else {
            for (uint32_t i = t; i < nq; i += kFuseB) Qg[i] = sk[i];
        }
        for (uint32_t e = 0; e < kFuseF; ++e)
            for (int b = 0; b < 8; ++b)
                if ((valid[e] >> b) & 1u) {
                    const Key v = morton_encode((b & 1) ? pc1[e][0] : pc0[e][0], (b & 2) ? pc1[e][1] : pc0[e][1],
                                                (b & 4) ? pc1[e][2] : pc0[e][2], lvl - 1, md);
                    if (stay) wl[npos++] = v;
                    else Wn[npos++] = v;
                }
        __syncthreads();
        PQ = Qg;
        pq_lds = stay;
        nprev = nq;
        l -= 1;
        nw = nwn;
        W = Wn;
        w_lds = stay;
        from_inpar = false;
        // levels without seeds: the next level starts from its input parents
        // only, with no split octants below it
        while (nw == 0 && l >= 1) {
            nprev = 0;
            if (l >= 2) {
                nw = s_loff[l] - s_loff[l - 1];
                W = inpar + s_loff[l - 1];
                w_lds = false;
                from_inpar = true;
            }
            l -= 1;
        }
        k += 1;
        if (l == 0 || nw > kFuseCap) break;
    }
    if (t == 0) {
        st->level = l;
        st->nw = nw;
        st->wsel = (k & 1) ? 0u : 1u;  // the buffer written last: (k - 1) & 1
        st->from_inpar = from_inpar ? 1u : 0u;
        st->nprev = nprev;
        st->qsel = (k & 1) ? 0u : 1u;
        st->used = used;
        st->levels = levels;
        st->seeds = seeds;
    }
}

}  // namespace

// ------------------------------------------------------------------ host ----
static void copy_keys(Key *dst, const Key *src, size_t n, hipStream_t s) {
    if (n) HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
}

DevBuf<Key> p2o_build(const double *d_xyz, size_t n, uint32_t md, uint32_t m, int curve, hipStream_t s,
                      P2OStats *stats, const double *glens) {
    SfcPacked tab = make_packed(curve);
    if (md == 0 || n <= m) {
        DevBuf<Key> root(1, s);
        Key r{0ull, 0ull};
        HIP_OK(hipMemcpyAsync(root.p, &r, sizeof(Key), hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
        if (stats) *stats = P2OStats{};
        return root;
    }
    DevBuf<Key> pk(n, s);
    // device counters: [0] distinct leaf keys Q, [1] textbook leaves nT,
    // [2] a point outside the domain, [3] overlapping leaves (internal error),
    // [4] a run of equal high words too long for k_sort_hi_runs
    DevBuf<uint32_t> ctr(5, s);
    HIP_OK(hipMemsetAsync(ctr.p, 0, 5 * sizeof(uint32_t), s));
    const double gx = glens ? glens[0] : 1.0, gy = glens ? glens[1] : 1.0, gz = glens ? glens[2] : 1.0;
    k_quantise<<<grid_for(n, B), B, 0, s>>>(d_xyz, n, md, gx, gy, gz, tab, pk.p, ctr.p + 2);
    HIP_OK(hipGetLastError());
    // Every key is at level md and bits 5-32 of the low word are zero by the
    // key layout: bits 32-127 decide the order.  The low word holds levels 22
    // and finer, which only order points that share their level-21 cell.  So
    // the keys are sorted on the high word alone (8 digits instead of 12) and
    // the short runs of equal high words on their low words; only a run longer
    // than kHiRunCap has the keys sorted on bits 32-127 and the steps redone.
    DevBuf<uint8_t> lev;
    DevBuf<Key> leafkey;
    DevBuf<uint32_t> NL, cnt, off;
    uint32_t hc[5];
    for (int begin_bit = 64;; begin_bit = 32) {
        radix_sort_keys(pk.p, nullptr, n, s, begin_bit, 128);
        if (begin_bit == 64) k_sort_hi_runs<<<grid_for(n, B), B, 0, s>>>(pk.p, n, ctr.p + 4);
        // textbook leaves
        lev.alloc(n - m, s);
        k_p2o_lev<<<grid_for(n - m, B), B, 0, s>>>(pk.p, n, m, md, lev.p);
        leafkey.alloc(n, s);
        k_p2o_leaf<<<grid_for(n, B), B, 0, s>>>(pk.p, n, m, lev.p, leafkey.p);
        lev.release();
        DevBuf<uint32_t> head(n, s), hpos(n, s);
        k_head<<<grid_for(n, B), B, 0, s>>>(leafkey.p, n, head.p);
        scan_exclusive_u32_async(head.p, hpos.p, n, ctr.p, s);
        // the distinct leaves stay counted on the device: the arrays below are
        // sized by the capacity n and the kernels read Q
        NL.alloc(n, s);
        k_compact_idx<<<grid_for(n, B), B, 0, s>>>(head.p, hpos.p, n, NL.p);
        head.release();
        hpos.release();
        cnt.alloc(n + 1, s);
        off.alloc(n + 1, s);
        k_gap_count<<<grid_for(n + 1, B), B, 0, s>>>(leafkey.p, NL.p, ctr.p, n, cnt.p, ctr.p + 3);
        scan_exclusive_u32_async(cnt.p, off.p, n + 1, ctr.p + 1, s);
        dgx_read_back(hc, ctr.p, sizeof hc, s);
        if (hc[2]) {
            dgx_set_error("points2Octree: coordinates must lie in [0,1)");
            throw DgxError();
        }
        if (!hc[4] || begin_bit == 32) break;
        HIP_OK(hipMemsetAsync(ctr.p + 3, 0, 2 * sizeof(uint32_t), s));
    }
    if (hc[3]) {
        dgx_set_error("points2Octree: internal error, non-empty leaves overlap");
        throw DgxError();
    }
    const uint64_t nT = hc[1];
    cnt.release();
    DevBuf<Key> T(nT, s);
    DevBuf<uint32_t> tb(nT, s), tc(nT, s);
    k_gap_emit<<<grid_for(n + 1, B), B, 0, s>>>(leafkey.p, NL.p, ctr.p, n, off.p, T.p, tb.p, tc.p);
    HIP_OK(hipGetLastError());
    leafkey.release();
    NL.release();
    off.release();
    // saturated-cell triggers
    DevBuf<uint32_t> flag(nT, s), fpos(nT, s);
    k_p2o_trigger<<<grid_for(nT, B), B, 0, s>>>(T.p, tb.p, tc.p, nT, n, m, md, flag.p);
    uint64_t ntrig = scan_exclusive_u32(flag.p, fpos.p, nT, s);
    P2OStats st{};
    st.textbook_leaves = nT;
    st.triggers = ntrig;
    if (ntrig == 0) {
        if (stats) *stats = st;
        return T;
    }
    DevBuf<uint32_t> trig(ntrig, s);
    k_compact_idx<<<grid_for(nT, B), B, 0, s>>>(flag.p, fpos.p, nT, trig.p);
    std::vector<uint32_t> htrig(ntrig);
    dgx_read_back(htrig.data(), trig.p, ntrig * 4, s);
    struct Seg { int src; size_t begin, count; };  // src 0: T, 1: replay buffer
    std::vector<Seg> segs;
    const uint32_t cap = 1u << 20;
    DevBuf<Key> rbuf(cap, s);
    DevBuf<ReplayResult> dres(1, s);
    size_t t_next = 0, rused = 0;
    for (uint32_t k : htrig) {
        if (k < t_next) continue;  // inside a replayed stretch
        segs.push_back({0, t_next, (size_t)k + 1 - t_next});
        k_p2o_replay<<<1, 64, 0, s>>>(pk.p, n, m, md, T.p, nT, tb.p, tc.p, k, rbuf.p + rused,
                                     (uint32_t)(cap - rused), dres.p);
        HIP_OK(hipGetLastError());
        ReplayResult hr;
        dgx_read_back(&hr, dres.p, sizeof(hr), s);
        if (hr.err) {
            dgx_set_error("points2Octree: cursor replay failed (code %u) after leaf %u", hr.err, k);
            throw DgxError();
        }
        segs.push_back({1, rused, hr.nout});
        rused += hr.nout;
        st.replayed_leaves += hr.nout;
        st.replays++;
        if (hr.end) { t_next = nT; break; }
        t_next = hr.next_t;  // first textbook leaf at or inside the resync octant
    }
    if (t_next < nT) segs.push_back({0, t_next, nT - t_next});
    size_t total = 0;
    for (auto &g : segs) total += g.count;
    DevBuf<Key> outk(total, s);
    size_t o = 0;
    for (auto &g : segs) {
        copy_keys(outk.p + o, (g.src ? rbuf.p : T.p) + g.begin, g.count, s);
        o += g.count;
    }
    HIP_OK(hipStreamSynchronize(s));
    st.leaves = total;
    if (stats) *stats = st;
    return outk;
}

DevBuf<Key> balance_build(const Key *in, size_t n, uint32_t md, int curve, hipStream_t s, BalStats *stats) {
    DevBuf<Key> out;
    BalStats st{};
    st.levels = 0;
    auto copy_input = [&]() {
        out.alloc(n, s);
        copy_keys(out.p, in, n, s);
        HIP_OK(hipStreamSynchronize(s));
        if (stats) *stats = st;
        return std::move(out);
    };
    if (n < 2) return copy_input();  // comboRipple: in.size() < 2 (balance.cpp:806-808)
    // level histogram of the inputs (also gives the level range)
    DevBuf<uint32_t> lh(32, s), cursor(32, s);
    HIP_OK(hipMemsetAsync(lh.p, 0, 128, s));
    k_level_hist<<<grid_for(n, B * LH_ITEMS), B, 0, s>>>(in, n, lh.p);
    uint32_t hl[32];
    dgx_read_back(hl, lh.p, 128, s);
    uint32_t lmin = 31, lmax = 0;
    for (uint32_t l = 0; l < 32; ++l)
        if (hl[l]) {
            lmin = std::min(lmin, l);
            lmax = std::max(lmax, l);
        }
    if (lmax - lmin < 2) return copy_input();  // TreeNode.cpp:929-947
    SfcPacked tab = make_packed(curve);
    // bucket the parents of the inputs by input level
    uint32_t loff[33];
    loff[0] = 0;
    for (int l = 0; l < 32; ++l) loff[l + 1] = loff[l] + hl[l];
    HIP_OK(hipMemcpyAsync(cursor.p, loff, 128, hipMemcpyHostToDevice, s));
    DevBuf<Key> inpar(n, s);
    k_bucket_parents<<<grid_for(n, B * LH_ITEMS), B, 0, s>>>(in, n, md, tab, cursor.p, inpar.p);
    HIP_OK(hipGetLastError());
    const uint32_t maxL = lmax;
    // Per level: sort + unique the seeds W (parents at level l-1), emit the
    // leaves of level l, generate the neighbour parents that seed level l-1.
    // Counts stay on the device (cnt[0] split octants, cnt[1] leaves, cnt[2]
    // neighbour parents); the host reads them once per level.
    DevBuf<uint32_t> cnt(4, s);
    uint32_t *d_nq = cnt.p, *d_nlv = cnt.p + 1, *d_nnb = cnt.p + 2;
    // single-pass compactions: look-back words and one tile counter per call
    // (epoch), zeroed together and reused until the counters run out
    DevBuf<uint64_t> lb_status;
    DevBuf<uint32_t> lb_ctr(128, s);
    HIP_OK(hipMemsetAsync(lb_ctr.p, 0, lb_ctr.n * sizeof(uint32_t), s));
    uint32_t lb_epoch = 0;
    uint32_t *fault = dgx_fault_word();
    DevBuf<Key> prevQ;  // split octants one level finer (sorted unique Morton)
    size_t nprev = 0;
    DevBuf<Key> W;      // seeds of the current level: input parents, then neighbour parents
    size_t nw = 0;
    DevBuf<Key> leaves(2 * n + 4096, s);
    size_t used = 0;
    // seeds of the finest level: input parents only
    nw = hl[maxL];
    W.alloc(nw, s);
    copy_keys(W.p, inpar.p + loff[maxL], nw, s);
    // buffers of the single-workgroup runs over small levels (k_bal_fused)
    DevBuf<Key> fw0, fw1, fq0, fq1;
    DevBuf<uint32_t> d_loff;
    DevBuf<FuseState> fst;
    uint32_t maxhl = 0;
    for (int l = 0; l < 32; ++l) maxhl = std::max(maxhl, hl[l]);
    uint32_t l = maxL;
    while (l >= 1) {
        if (nw == 0) {
            prevQ.release();
            nprev = 0;
            if (l >= 2) {
                nw = hl[l - 1];
                W.alloc(nw, s);
                copy_keys(W.p, inpar.p + loff[l - 1], nw, s);
            }
            --l;
            continue;
        }
        if (nw <= kFuseCap) {
            // run levels l, l-1, ... in one workgroup while their seeds stay small
            if (!fst.p) {
                const size_t wcap = (size_t)maxhl + 8 * (size_t)kFuseCap;
                fw0.alloc(wcap, s);
                fw1.alloc(wcap, s);
                fq0.alloc(kFuseCap, s);
                fq1.alloc(kFuseCap, s);
                d_loff.alloc(33, s);
                HIP_OK(hipMemcpyAsync(d_loff.p, loff, 33 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
                fst.alloc(1, s);
            }
            const size_t need = used + (size_t)8 * kFuseCap * l;  // at most 8 leaves per split octant
            if (leaves.n < need) {
                DevBuf<Key> g(std::max(2 * leaves.n, need), s);
                copy_keys(g.p, leaves.p, used, s);
                leaves = std::move(g);
            }
            k_bal_fused<<<1, kFuseB, 0, s>>>(l, W.p, (uint32_t)nw, prevQ.p, (uint32_t)nprev, inpar.p, d_loff.p, md,
                                             tab, leaves.p, (uint32_t)used, fw0.p, fw1.p, fq0.p, fq1.p, fst.p);
            HIP_OK(hipGetLastError());
            FuseState h;
            dgx_read_back(&h, fst.p, sizeof h, s);
            used = h.used;
            st.levels += h.levels;
            st.seeds += h.seeds;
            if (h.level == 0) break;
            l = h.level;
            nw = h.nw;
            nprev = h.nprev;
            prevQ.alloc(nprev, s);
            copy_keys(prevQ.p, h.qsel ? fq1.p : fq0.p, nprev, s);
            W.alloc(nw, s);
            copy_keys(W.p, h.from_inpar ? inpar.p + loff[l] : (h.wsel ? fw1.p : fw0.p), nw, s);
            continue;
        }
        // parents live at level lv = l-1: only digits 1..lv vary
        const uint32_t lv = l - 1;
        const int begin_bit = lv <= 21 ? 127 - 3 * (int)lv : 63 - 3 * ((int)lv - 21);
        radix_sort_keys(W.p, nullptr, nw, s, begin_bit, 127);
        const unsigned fs_grid = (unsigned)((nw + FS_TILE - 1) / FS_TILE);
        const unsigned le_grid = (unsigned)((nw + LE_TILE - 1) / LE_TILE);  // at least fs_grid
        if (le_grid > lb_status.n) {  // look-back words for the largest level so far, zeroed
            lb_status.alloc(std::max<size_t>(le_grid, 2 * lb_status.n), s);
            HIP_OK(hipMemsetAsync(lb_status.p, 0, lb_status.n * sizeof(uint64_t), s));
        }
        if (lb_epoch + 2 >= lb_ctr.n) {  // a zeroed tile counter per call
            HIP_OK(hipMemsetAsync(lb_ctr.p, 0, lb_ctr.n * sizeof(uint32_t), s));
            HIP_OK(hipMemsetAsync(lb_status.p, 0, lb_status.n * sizeof(uint64_t), s));
            lb_epoch = 0;
        }
        DevBuf<Key> Q(nw, s);  // capacity; *d_nq used
        DevBuf<uint32_t> cmask(nw, s);
        k_unique_compact<<<fs_grid, FS_B, 0, s>>>(W.p, nw, Q.p, cmask.p, d_nq, d_nnb, lb_status.p, lb_ctr.p,
                                                  ++lb_epoch, fault);
        // leaves at level l: children of Q not split themselves
        if (nprev) k_mark_split_children<<<grid_for(nprev, B), B, 0, s>>>(prevQ.p, nprev, Q.p, d_nq, cmask.p);
        if (leaves.n < used + 8 * nw) {  // grow: at most 8 leaves per split octant
            DevBuf<Key> g(std::max(2 * leaves.n, used + 8 * nw), s);
            copy_keys(g.p, leaves.p, used, s);
            leaves = std::move(g);
        }
        k_leaf_emit_scan<<<le_grid, LE_B, 0, s>>>(Q.p, d_nq, cmask.p, md, tab, leaves.p + used, d_nlv, lb_status.p,
                                                  lb_ctr.p, ++lb_epoch, fault);
        HIP_OK(hipGetLastError());
        // neighbours of the split octants seed the next coarser level
        DevBuf<Key> Wn;
        const size_t nin_next = l >= 2 ? hl[l - 1] : 0;
        if (l >= 2) {
            Wn.alloc(nin_next + 8 * nw, s);
            k_bal_nbr_parents<<<grid_for(std::max(nw, nin_next), B * NB_ITEMS), B, 0, s>>>(
                Q.p, d_nq, l - 1, md, Wn.p, d_nnb, inpar.p + loff[l - 1], nin_next);
            HIP_OK(hipGetLastError());
        }
        uint32_t hc[3];
        dgx_read_back(hc, cnt.p, 12, s);
        used += hc[1];
        st.levels++;
        st.seeds += nw;
        prevQ = std::move(Q);
        nprev = hc[0];
        W = std::move(Wn);
        nw = nin_next + hc[2];
        --l;
    }
    W.release();
    prevQ.release();
    const size_t total = used;
    // Distinct leaves are never ancestor and descendant, so a digit at or above
    // the coarser level tells them apart: the level byte needs no pass.  Only
    // leaves finer than level 21 that share a level-21 cell have equal high
    // words, so the high-word sort and k_sort_hi_runs order them, and bits
    // 8-127 are sorted in full only when such a cell holds too many leaves.
    DevBuf<uint32_t> longrun(1, s);
    HIP_OK(hipMemsetAsync(longrun.p, 0, sizeof(uint32_t), s));
    radix_sort_keys(leaves.p, nullptr, total, s, 64, 128);
    if (total) k_sort_hi_runs<<<grid_for(total, B), B, 0, s>>>(leaves.p, total, longrun.p);
    HIP_OK(hipGetLastError());
    uint32_t hlong = 0;
    dgx_read_back(&hlong, longrun.p, sizeof hlong, s);
    if (hlong) {
        radix_sort_keys(leaves.p, nullptr, total, s, 8, 128);
        HIP_OK(hipStreamSynchronize(s));
    }
    out = std::move(leaves);
    out.n = total;
    st.leaves = total;
    if (stats) *stats = st;
    return out;
}

void export_treenodes(const Key *k, size_t n, uint32_t md, int curve, uint32_t *d_tn, hipStream_t s) {
    if (!n) return;
    k_export_tn<<<grid_for(n, B), B, 0, s>>>(k, n, md, make_packed(curve), d_tn);
    HIP_OK(hipGetLastError());
}

DevBuf<Key> import_treenodes(const uint32_t *d_tn, size_t n, uint32_t md, int curve, bool sort, hipStream_t s) {
    DevBuf<Key> k(n, s);
    DevBuf<uint32_t> bad(1, s);
    HIP_OK(hipMemsetAsync(bad.p, 0, 4, s));
    if (n) k_import_tn<<<grid_for(n, B), B, 0, s>>>(d_tn, n, md, make_packed(curve), k.p, bad.p);
    HIP_OK(hipGetLastError());
    uint32_t hb = 0;
    dgx_read_back(&hb, bad.p, 4, s);
    if (hb) {
        dgx_set_error("TreeNode array: anchor not aligned to its level or outside the domain");
        throw DgxError();
    }
    if (sort) radix_sort_keys(k.p, nullptr, n, s, 0, 128);
    return k;
}
