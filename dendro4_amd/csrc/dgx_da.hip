// dgx_da.hip -- ot::DA construction on one rank (MPI_COMM_SELF) for gfx950.
//
// Follows the reference factory (src/oda/private/odaFactory.cpp:102-923):
//   Part1 (:150-205): embed every octant one level deeper in a 2^(D+1) domain and
//         add the positive-boundary pseudo-octants (octFunctions.cpp:681-770),
//         then re-sort on the curve.  The original domain is Morton child 0 of
//         the embedded root, so its octants precede all pseudo-octants.
//   Part2 (:207-220): NODE flags (flagNodesType3, octFunctions.cpp:1556-2168):
//         an anchor is hanging iff the parent of one of the -x/-y/-z neighbours
//         selected by the child number is itself an octant.
//   Part3: one rank has no ghosts, so the local buffer is [elements | pseudo-
//         octants] and buildNodeList (odaBuildNlist.cpp:47-934, 3010-3095) is
//         8 searches per element: the vertex's owner (maxLowerBound at the
//         vertex cell) if it is a NODE anchored there, else the owner of the
//         parent's vertex, else the forward ghost scan.
// Node vectors are compacted in local-buffer order over NODE octants
// (DA::vecGetBuffer, oda.cpp:863-882).
#include "dgx_da.h"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace {

constexpr int B = 256;
constexpr uint8_t F_BOUNDARY = 64, F_NODE = 128;

#define LOAD_TABLES(tab)                                                 \
    __shared__ uint8_t s_rc[192], s_mc[192];                             \
    for (int i_ = threadIdx.x; i_ < 192; i_ += blockDim.x) {             \
        s_rc[i_] = (tab).rc[i_];                                          \
        s_mc[i_] = (tab).mc[i_];                                          \
    }                                                                    \
    __syncthreads();

__device__ inline uint32_t pos_bdy_flags(uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    // TreeNode::isBoundaryOctant(POSITIVE) (TreeNode.cpp:247-269)
    uint32_t len = 1u << (md - lev), blen = (1u << md) - len, f = 0;
    if (x == blen) f |= 1u;
    if (y == blen) f |= 2u;
    if (z == blen) f |= 4u;
    return f;
}
__device__ inline uint32_t n_pseudo(uint32_t f) {
    // +x, +y, +z, +x+y, +x+z, +y+z, +x+y+z (octFunctions.cpp:702-757)
    uint32_t c = __popc(f);
    return c == 3 ? 7u : (c == 2 ? 3u : c);
}

__global__ void __launch_bounds__(B) k_da_embed(const Key *in, size_t n, uint32_t md0, SfcPacked tab, Key *dkey,
                                                uint32_t *nbdy) {
    LOAD_TABLES(tab)
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z, l;
    dev_decode(s_mc, in[i], md0, x, y, z, l);
    dkey[i] = dev_encode(s_rc, x, y, z, l + 1, md0 + 1);
    nbdy[i] = n_pseudo(pos_bdy_flags(x, y, z, l, md0));
}

__global__ void __launch_bounds__(B) k_da_bdy_emit(const Key *in, size_t n, uint32_t md0, SfcPacked tab,
                                                   const uint32_t *off, Key *bkey) {
    LOAD_TABLES(tab)
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z, l;
    dev_decode(s_mc, in[i], md0, x, y, z, l);
    uint32_t f = pos_bdy_flags(x, y, z, l, md0);
    if (!f) return;
    uint32_t o = off[i], D = 1u << md0, md = md0 + 1, L = l + 1;
    if (f & 1u) bkey[o++] = dev_encode(s_rc, D, y, z, L, md);
    if (f & 2u) bkey[o++] = dev_encode(s_rc, x, D, z, L, md);
    if (f & 4u) bkey[o++] = dev_encode(s_rc, x, y, D, L, md);
    if ((f & 3u) == 3u) bkey[o++] = dev_encode(s_rc, D, D, z, L, md);
    if ((f & 5u) == 5u) bkey[o++] = dev_encode(s_rc, D, y, D, L, md);
    if ((f & 6u) == 6u) bkey[o++] = dev_encode(s_rc, x, D, D, L, md);
    if (f == 7u) bkey[o++] = dev_encode(s_rc, D, D, D, L, md);
}

__global__ void __launch_bounds__(B) k_da_decode(const Key *k, size_t n, size_t nelem, uint32_t md, SfcPacked tab,
                                                 uint32_t *xs, uint32_t *ys, uint32_t *zs, uint8_t *flags) {
    LOAD_TABLES(tab)
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z, l;
    dev_decode(s_mc, k[i], md, x, y, z, l);
    xs[i] = x;
    ys[i] = y;
    zs[i] = z;
    flags[i] = (uint8_t)(l | (i >= nelem ? F_BOUNDARY : 0));
}

// Search ranges of the sorted local keys by their cell at level 6 (the high
// word's bits 62..45): every key whose first 18 bits are c lies in
// [cstart[c], cstart[c + 1]), keys before cstart[c] are smaller than any key
// of cell c and keys from cstart[c + 1] on larger.  A search starts from its
// key's cell instead of the whole array and finds the same index.
constexpr uint32_t kCellBits = 18, kCellShift = 63 - kCellBits;
__device__ inline uint32_t cell_of(const Key &k) { return (uint32_t)(k.hi >> kCellShift); }
__global__ void k_cell_starts(const Key *k, size_t n, uint32_t *cstart) {
    const uint32_t c = blockIdx.x * B + threadIdx.x;
    if (c > (1u << kCellBits)) return;
    if (c == (1u << kCellBits)) {  // past every key (bit 63 of the high word is 0)
        cstart[c] = (uint32_t)n;
        return;
    }
    const uint64_t h = (uint64_t)c << kCellShift;  // first key whose high word is at least h
    size_t lo = 0, hi = n;
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (k[mid].hi < h) lo = mid + 1;
        else hi = mid;
    }
    cstart[c] = (uint32_t)lo;
}
__device__ inline bool bsearch_exact(const Key *a, size_t n, const uint32_t *cstart, const Key &k) {
    const uint32_t c = cell_of(k);
    size_t lo = cstart[c], hi = cstart[c + 1];
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_lt(a[mid], k)) lo = mid + 1; else hi = mid;
    }
    return lo < n && key_eq(a[lo], k);
}
// seq::maxLowerBound (seqUtils.tcc:59-100): greatest a[i] <= k.
__device__ inline bool max_lower_bound(const Key *a, size_t n, const uint32_t *cstart, const Key &k, size_t &idx) {
    const uint32_t c = cell_of(k);
    size_t lo = cstart[c], hi = cstart[c + 1];
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_le(a[mid], k)) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) { idx = 0; return false; }
    idx = lo - 1;
    return true;
}

// The probes' keys come from the octant's own curve path (path_walk once,
// then key_from_path: only the levels below the first one at which the probe
// leaves the octant's ancestors are walked) instead of a full encode each.
__global__ void __launch_bounds__(B) k_da_flag_nodes(const Key *k, size_t n, const uint32_t *__restrict__ cstart,
                                                     uint32_t md, SfcPacked tab, const uint32_t *xs,
                                                     const uint32_t *ys, const uint32_t *zs, uint8_t *flags) {
    LOAD_TABLES(tab)
    __shared__ uint8_t s_rot[kTailCols * B];  // rotation states of each thread's path, levels 1..32
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    uint32_t lev = flags[i] & 31u;
    bool hanging = false;
    uint8_t *rots = s_rot + threadIdx.x;
    if (lev != 1) {
        uint32_t X = xs[i], Y = ys[i], Z = zs[i], s = 1u << (md - lev);
        uint32_t cn = (((Z >> (md - lev)) & 1u) << 2) | (((Y >> (md - lev)) & 1u) << 1) | ((X >> (md - lev)) & 1u);
        const Key kref = (cn != 0 && cn != 7) ? path_walk(s_rc, X, Y, Z, lev, md, rots, B) : Key{0ull, 0ull};
        // parent of the same-level neighbour at (a,b,c), as a key at level lev-1
        auto probe = [&](uint32_t a, uint32_t b, uint32_t c) {
            uint32_t ps = md - lev + 1;
            Key pk = key_from_path(s_rc, kref, rots, B, X, Y, Z, (a >> ps) << ps, (b >> ps) << ps, (c >> ps) << ps,
                                   lev - 1, lev - 1, md, nullptr);
            if (bsearch_exact(k, n, cstart, pk)) hanging = true;
        };
        switch (cn) {
            case 1: if (Y) probe(X, Y - s, Z); if (Z) probe(X, Y, Z - s); if (Y && Z) probe(X, Y - s, Z - s); break;
            case 2: if (X) probe(X - s, Y, Z); if (Z) probe(X, Y, Z - s); if (X && Z) probe(X - s, Y, Z - s); break;
            case 3: if (Z) probe(X, Y, Z - s); break;
            case 4: if (X) probe(X - s, Y, Z); if (Y) probe(X, Y - s, Z); if (X && Y) probe(X - s, Y - s, Z); break;
            case 5: if (Y) probe(X, Y - s, Z); break;
            case 6: if (X) probe(X - s, Y, Z); break;
            default: break;
        }
    }
    if (!hanging) flags[i] |= F_NODE;
}

// buildNodeList for one element (odaBuildNlist.cpp:531-934 on one rank).
// The vertex keys come from the element's curve path (path_walk once, then
// key_from_path with the zero-digit tails of the level-md keys in `tail`,
// k_zero_tail) instead of a 31-level encode each.
__global__ void __launch_bounds__(B) k_da_nlist(const Key *k, size_t n, const uint32_t *__restrict__ cstart,
                                                size_t nelem, uint32_t md, SfcPacked tab,
                                                const uint32_t *xs, const uint32_t *ys, const uint32_t *zs,
                                                const uint8_t *flags, const Key *__restrict__ tail, uint32_t *nlist,
                                                uint8_t *mask,
                                                uint32_t *counters /* [0]=unresolved [1]=ghost scans */) {
    LOAD_TABLES(tab)
    __shared__ uint8_t s_rot[kTailCols * B];  // rotation states of each thread's path, levels 1..32
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i >= nelem) return;
    const uint32_t d = flags[i] & 31u, x = xs[i], y = ys[i], z = zs[i];
    uint8_t *rots = s_rot + threadIdx.x;
    const Key kref = path_walk(s_rc, x, y, z, d, md, rots, B);
    const bool inode = flags[i] & F_NODE;
    const uint32_t ps = md - d + 1, parX = (x >> ps) << ps, parY = (y >> ps) << ps, parZ = (z >> ps) << ps;
    const uint32_t sz = 1u << (md - d);
    uint32_t nl[8];
    for (int j = 0; j < 8; ++j) {
        const uint32_t bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
        const uint32_t vx = x + bx * sz, vy = y + by * sz, vz = z + bz * sz;
        const uint32_t wx = parX + bx * (sz << 1), wy = parY + by * (sz << 1), wz = parZ + bz * (sz << 1);
        size_t idx = 0;
        bool found;
        if (j == 0) {
            uint32_t r = (uint32_t)i;
            if (!inode) {
                Key pk = key_from_path(s_rc, kref, rots, B, x, y, z, wx, wy, wz, d - 1, md, md, tail);
                found = max_lower_bound(k, n, cstart, pk, idx);
                if (found && xs[idx] == wx && ys[idx] == wy && zs[idx] == wz && (flags[idx] & F_NODE)) r = (uint32_t)idx;
            }
            nl[0] = r;
            continue;
        }
        Key sk = key_from_path(s_rc, kref, rots, B, x, y, z, vx, vy, vz, d, md, md, tail);
        uint32_t sx = vx, sy = vy, szz = vz, lastLevel = d;
        found = max_lower_bound(k, n, cstart, sk, idx);
        if (found && !(key_is_ancestor(k[idx], sk) || key_eq(k[idx], sk))) found = false;
        bool secondary = true;
        if (found && xs[idx] == vx && ys[idx] == vy && zs[idx] == vz && (flags[idx] & F_NODE)) {
            nl[j] = (uint32_t)idx;
            secondary = false;
        }
        if (!secondary) continue;
        sx = wx; sy = wy; szz = wz; lastLevel = d - 1;
        Key pk = key_from_path(s_rc, kref, rots, B, x, y, z, wx, wy, wz, d - 1, md, md, tail);
        found = max_lower_bound(k, n, cstart, pk, idx);
        if (found && xs[idx] == wx && ys[idx] == wy && zs[idx] == wz && (flags[idx] & F_NODE)) {
            nl[j] = (uint32_t)idx;
            continue;
        }
        // forward ghost scan (odaBuildNlist.cpp:566-600)
        atomicAdd(&counters[1], 1u);
        Key nk = dev_encode(s_rc, sx, sy, szz, lastLevel, md);
        Key dld = nk;  // last cell of nk
        for (uint32_t t = lastLevel + 1; t <= md; ++t) key_set_digit(dld, t, 7);
        dld.lo = (dld.lo & ~31ull) | md;
        uint32_t res = (uint32_t)n;
        for (size_t kk = 1; idx + kk < n; ++kk) {
            const Key &c = k[idx + kk];
            if (key_eq(key_parent(c), nk)) {
                if (flags[idx + kk] & F_NODE) continue;
                res = (uint32_t)(idx + kk);
                break;
            }
            if (key_lt(dld, c)) break;
        }
        nl[j] = res;
    }
    uint8_t m = 0;
    for (int j = 0; j < 8; ++j) {
        uint32_t r = nl[j];
        nlist[8 * i + j] = r;
        if (r >= n) {
            atomicAdd(&counters[0], 1u);
            continue;
        }
        uint32_t _x = xs[r], _y = ys[r], _z = zs[r];
        if (!(flags[r] & F_NODE)) {
            uint32_t s2 = md - (flags[r] & 31u) + 1;
            _x = (_x >> s2) << s2; _y = (_y >> s2) << s2; _z = (_z >> s2) << s2;
        }
        const uint32_t bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
        if (j == 0) {
            if (!inode) m |= 1;
        } else if (x + bx * sz != _x || y + by * sz != _y || z + bz * sz != _z) {
            m |= (uint8_t)(1u << j);
        }
    }
    mask[i] = m;
}

// zero_tail table of the level-md keys: tail[r * kTailCols + s], r < 24.
__global__ void k_zero_tail(SfcPacked tab, uint32_t md, Key *tail) {
    for (uint32_t e = threadIdx.x; e < 24 * kTailCols; e += blockDim.x)
        tail[e] = zero_tail(tab.rc, e / kTailCols, e % kTailCols, md);
}

__global__ void k_da_mark_nodes(const uint32_t *nlist, size_t nelem, size_t n, uint8_t *isnode) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * nelem) return;
    uint32_t r = nlist[t];
    if (r < n) isnode[r] = 1;
}

// Grid-stride over a capped grid; the two counts are summed per workgroup and
// added with one atomic each (one atomic per wave on the same two words cost
// ~5 ns each, serialised: 2.4 ms on a 30M-octant local buffer).
__global__ void __launch_bounds__(B) k_da_apply_nodes(const uint8_t *isnode, size_t n, size_t nelem, uint8_t *flags,
                                                      uint32_t *nodeflag,
                                                      uint32_t *counts /* [0] elem nodes [1] bnd nodes */) {
    __shared__ uint32_t s_cnt[2];
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t ce = 0, cb = 0;
    for (size_t i = (size_t)blockIdx.x * B + threadIdx.x; i < n; i += (size_t)gridDim.x * B) {
        uint8_t f = flags[i];
        if (isnode[i]) {
            if (!(f & F_NODE)) flags[i] = f = (uint8_t)(f | F_NODE);
            if (i < nelem) ++ce;
            else ++cb;
        }
        nodeflag[i] = (f & F_NODE) ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ce += (uint32_t)__shfl_xor(ce, o, 64);
        cb += (uint32_t)__shfl_xor(cb, o, 64);
    }
    if (__lane_id() == 0) {
        if (ce) atomicAdd(&s_cnt[0], ce);
        if (cb) atomicAdd(&s_cnt[1], cb);
    }
    __syncthreads();
    if (threadIdx.x < 2 && s_cnt[threadIdx.x]) atomicAdd(&counts[threadIdx.x], s_cnt[threadIdx.x]);
}

__global__ void k_da_l2n(const uint32_t *nodeflag, const uint32_t *pos, size_t n, uint32_t *l2n, uint32_t *n2l) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    l2n[i] = nodeflag[i] ? pos[i] : 0xffffffffu;
    if (nodeflag[i]) n2l[pos[i]] = (uint32_t)i;
}
// Nodes on the boundary of the unit cube (anchor coordinate 0 or 2^(md-1)).
__global__ void k_da_bdy_nodes(const uint32_t *n2l, uint32_t nn, const uint32_t *xs, const uint32_t *ys,
                               const uint32_t *zs, uint32_t md, uint8_t *bdy, uint32_t *count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool on = false;
    if (i < nn) {
        const uint32_t D = 1u << (md - 1), l = n2l[i], x = xs[l], y = ys[l], z = zs[l];
        on = x == 0 || y == 0 || z == 0 || x == D || y == D || z == D;
        bdy[i] = on ? 1 : 0;
    }
    uint32_t c = __popcll(__ballot(on));
    if (__lane_id() == 0 && c) atomicAdd(count, c);
}

// Element view for the matvec: SoA compact node ids and packed meta
// (hanging mask | child number << 8 | original level << 11).
__global__ void k_da_elem_view(const uint32_t *nlist, const uint8_t *mask, const uint32_t *l2n, const uint32_t *xs,
                               const uint32_t *ys, const uint32_t *zs, const uint8_t *flags, size_t nelem,
                               uint32_t md, uint32_t *conn, uint16_t *meta) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    for (int j = 0; j < 8; ++j) conn[(size_t)j * nelem + e] = l2n[nlist[8 * e + j]];
    uint32_t lev = flags[e] & 31u, b = md - lev;
    uint32_t cn = (((zs[e] >> b) & 1u) << 2) | (((ys[e] >> b) & 1u) << 1) | ((xs[e] >> b) & 1u);
    meta[e] = (uint16_t)(mask[e] | (cn << 8) | ((lev - 1) << 11));
}

}  // namespace

DaDevice::~DaDevice() {}

std::unique_ptr<DaDevice> da_build(const Key *oct, size_t n, uint32_t md0, int curve, hipStream_t s, bool colour) {
    if (n <= 8) {  // blockPartStage1 asserts nodes.size() > 2^dim (blockPart.cpp:583)
        dgx_set_error("ot::DA needs more than 8 octants (got %zu)", n);
        throw DgxError();
    }
    std::unique_ptr<DaDevice> da(new DaDevice());
    da->md = md0 + 1;
    da->curve = curve;
    const uint32_t md = md0 + 1;
    SfcPacked tab = make_packed(curve);
    // Part1
    DevBuf<Key> dkey(n, s);
    DevBuf<uint32_t> nb(n, s), nboff(n, s);
    k_da_embed<<<grid_for(n, B), B, 0, s>>>(oct, n, md0, tab, dkey.p, nb.p);
    HIP_OK(hipGetLastError());
    uint64_t nbd = scan_exclusive_u32(nb.p, nboff.p, n, s);
    const size_t N = n + nbd;
    da->key.alloc(N, s);
    if (nbd) k_da_bdy_emit<<<grid_for(n, B), B, 0, s>>>(oct, n, md0, tab, nboff.p, da->key.p + n);
    HIP_OK(hipMemcpyAsync(da->key.p, dkey.p, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
    dkey.release();
    nb.release();
    nboff.release();
    radix_sort_keys(da->key.p, nullptr, n, s, 0, 128);
    if (nbd) radix_sort_keys(da->key.p + n, nullptr, nbd, s, 0, 128);
    da->n_local = N;
    da->n_elem = n;
    da->x.alloc(N, s);
    da->y.alloc(N, s);
    da->z.alloc(N, s);
    da->flags.alloc(N, s);
    k_da_decode<<<grid_for(N, B), B, 0, s>>>(da->key.p, N, n, md, tab, da->x.p, da->y.p, da->z.p, da->flags.p);
    // Part2
    DevBuf<uint32_t> cstart((1u << kCellBits) + 1, s);
    k_cell_starts<<<grid_for((1u << kCellBits) + 1, B), B, 0, s>>>(da->key.p, N, cstart.p);
    k_da_flag_nodes<<<grid_for(N, B), B, 0, s>>>(da->key.p, N, cstart.p, md, tab, da->x.p, da->y.p, da->z.p,
                                                 da->flags.p);
    HIP_OK(hipGetLastError());
    // Part3: node list
    da->nlist.alloc(8 * n, s);
    da->mask.alloc(n, s);
    DevBuf<uint32_t> ctr(4, s);
    HIP_OK(hipMemsetAsync(ctr.p, 0, 16, s));
    DevBuf<Key> tail(24 * kTailCols, s);
    k_zero_tail<<<1, B, 0, s>>>(tab, md, tail.p);
    k_da_nlist<<<grid_for(n, B), B, 0, s>>>(da->key.p, N, cstart.p, n, md, tab, da->x.p, da->y.p, da->z.p,
                                           da->flags.p, tail.p, da->nlist.p, da->mask.p, ctr.p);
    HIP_OK(hipGetLastError());
    DevBuf<uint8_t> isnode(N, s);
    HIP_OK(hipMemsetAsync(isnode.p, 0, N, s));
    k_da_mark_nodes<<<grid_for(8 * n, B), B, 0, s>>>(da->nlist.p, n, N, isnode.p);
    DevBuf<uint32_t> nodeflag(N, s), pos(N, s);
    k_da_apply_nodes<<<(unsigned)std::min<size_t>(grid_for(N, B), 2048), B, 0, s>>>(isnode.p, N, n, da->flags.p,
                                                                                  nodeflag.p, ctr.p + 2);
    uint64_t nn = scan_exclusive_u32(nodeflag.p, pos.p, N, s);
    da->l2n.alloc(N, s);
    da->n2l.alloc(nn, s);
    k_da_l2n<<<grid_for(N, B), B, 0, s>>>(nodeflag.p, pos.p, N, da->l2n.p, da->n2l.p);
    da->bdy_node.alloc(nn, s);
    DevBuf<uint32_t> bcount(1, s);
    HIP_OK(hipMemsetAsync(bcount.p, 0, 4, s));
    k_da_bdy_nodes<<<grid_for(nn, B), B, 0, s>>>(da->n2l.p, (uint32_t)nn, da->x.p, da->y.p, da->z.p, md,
                                                da->bdy_node.p, bcount.p);
    dgx_read_back(&da->n_bdy_nodes, bcount.p, 4, s);
    uint32_t hc[4];
    dgx_read_back(hc, ctr.p, 16, s);
    da->n_unresolved = hc[0];
    da->n_ghost_scans = hc[1];
    da->node_size = hc[2];
    da->bnd_node_size = hc[3];
    da->n_nodes = (uint32_t)nn;
    if (hc[0]) {
        dgx_set_error("ot::DA: %u node-list entries unresolved on one rank", hc[0]);
        throw DgxError();
    }
    if ((uint64_t)hc[2] + hc[3] != nn) {
        dgx_set_error("ot::DA: %llu NODE octants but %u referenced nodes", (unsigned long long)nn, hc[2] + hc[3]);
        throw DgxError();
    }
    // matvec view
    da->conn.alloc(8 * n, s);
    da->meta.alloc(n, s);
    k_da_elem_view<<<grid_for(n, B), B, 0, s>>>(da->nlist.p, da->mask.p, da->l2n.p, da->x.p, da->y.p, da->z.p,
                                               da->flags.p, n, md, da->conn.p, da->meta.p);
    HIP_OK(hipGetLastError());
    da_build_blocks(*da, s, colour);
    HIP_OK(hipStreamSynchronize(s));
    return da;
}

// ---------------------------------------------------------------------------
// Matvec block view.
// ---------------------------------------------------------------------------
namespace {

// (block, node) pair keys for every element vertex.
__global__ void k_blk_pairs(const uint32_t *conn, size_t ne, uint32_t eb, Key *pk) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * ne) return;
    size_t e = t % ne;
    pk[t] = Key{(uint64_t)(e / eb), (uint64_t)conn[t]};
}
__global__ void k_blk_head(const Key *pk, size_t n, uint32_t *head, uint32_t *bstart) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || !key_eq(pk[i], pk[i - 1])) ? 1u : 0u;
    if (i == 0 || pk[i].hi != pk[i - 1].hi) bstart[pk[i].hi] = (uint32_t)i;  // first raw index of each block
}
__global__ void k_blk_compact(const Key *pk, const uint32_t *head, const uint32_t *pos, size_t n, uint32_t *bnode,
                              uint32_t *pair_block, uint32_t *ncount) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    uint32_t p = pos[i];
    bnode[p] = (uint32_t)pk[i].lo;
    pair_block[p] = (uint32_t)pk[i].hi;
    atomicAdd(&ncount[pk[i].lo], 1u);
}
__global__ void k_blk_offsets(const uint32_t *bstart, const uint32_t *pos, uint32_t nblk, uint64_t npairs,
                              uint32_t *boff) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblk) return;
    boff[b] = (b == nblk) ? (uint32_t)npairs : pos[bstart[b]];
}
// Local index of every element vertex inside its block's sorted node list,
// stored in the child frame (slot j ^ c for vertex j of child c).
// Slot of rank r among n: ranks r, r+1, ..., r+7 (one sibling group in SFC
// order) go to 8 different rows of a row-major 8-row layout, so the 64 lanes of
// one instruction hold elements from 64 different sibling groups, which in the
// child frame touch 64 different nodes at each vertex (no LDS address conflicts).
__device__ __forceinline__ uint32_t spread8(uint32_t r, uint32_t n) {
    const uint32_t row = r & 7u, col = r >> 3;
    uint32_t base = 0;
    for (uint32_t i = 0; i < row; ++i) base += (n - i + 7u) >> 3;  // length of row i
    return base + col;
}
// Order of the elements inside each block for the matvec: elements without
// hanging vertices first, then the hanging ones, so that most waves are
// uniform; inside each class the SFC order is spread by spread8.
// epos[e] = the element's slot.
__global__ void __launch_bounds__(B) k_blk_order(const uint16_t *meta, size_t ne, uint32_t eb, uint32_t *epos) {
    __shared__ uint32_t wf[B / 64], wh[B / 64];
    __shared__ uint32_t total_flat, carry_f, carry_h;
    const size_t e0 = (size_t)blockIdx.x * eb;
    const uint32_t nb = (uint32_t)min((size_t)eb, ne - e0);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        total_flat = 0;
        carry_f = 0;
        carry_h = 0;
    }
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t k = threadIdx.x; k < nb; k += B) c += (meta[e0 + k] & 255u) == 0 ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) atomicAdd(&total_flat, c);
    __syncthreads();
    const uint32_t nflat = total_flat;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t k0 = 0; k0 < nb; k0 += B) {
        const uint32_t k = k0 + threadIdx.x;
        const bool valid = k < nb;
        const bool hang = valid && (meta[e0 + k] & 255u) != 0;
        const bool flat = valid && !hang;
        const uint64_t bf = __ballot(flat), bh = __ballot(hang);
        if (lane == 0) {
            wf[wid] = (uint32_t)__popcll(bf);
            wh[wid] = (uint32_t)__popcll(bh);
        }
        __syncthreads();
        uint32_t pf = carry_f, ph = carry_h;
        for (int w = 0; w < wid; ++w) {
            pf += wf[w];
            ph += wh[w];
        }
        if (flat) epos[e0 + k] = spread8(pf + (uint32_t)__popcll(bf & lt), nflat);
        if (hang) epos[e0 + k] = nflat + spread8(ph + (uint32_t)__popcll(bh & lt), nb - nflat);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 0; w < B / 64; ++w) {
                carry_f += wf[w];
                carry_h += wh[w];
            }
        __syncthreads();
    }
}
// lconn[(b*eb + k)*8 + j] (k = the element's slot in block b): byte offset
// (8 x local index) of vertex j ^ c of the element (child number c) in the
// block's node list; one element's 8 entries are one 16-byte load.
__global__ void k_blk_lconn(const uint32_t *conn, const uint16_t *meta, size_t ne, uint32_t eb, const uint32_t *boff,
                            const uint32_t *bnode, const uint32_t *epos, uint16_t *lconn) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * ne) return;
    size_t e = t % ne, j = t / ne;
    uint32_t b = (uint32_t)(e / eb), g = conn[t], c = (meta[e] >> 8) & 7u;
    uint32_t lo = boff[b], hi = boff[b + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (bnode[mid] < g) lo = mid + 1; else hi = mid;
    }
    lconn[((size_t)b * eb + epos[e]) * 8 + (j ^ c)] = (uint16_t)((lo - boff[b]) * 8u);
}
// One workgroup per element block of at most kBlkNodesE elements: the
// block's 8 x nb vertex node ids are sorted in LDS (bitonic; stages with
// partner distance <= 64 stay inside one wave's elements and synchronise that
// wave only), made unique, and written to ulist[b * 8 * eb ...] with their
// count in ucount[b]; every vertex's local index (lower bound in the unique
// list) goes to lconn in the child frame, as k_blk_lconn does.
constexpr uint32_t kBlkNodesE = 512, kBlkNodesB = 256;
__device__ __forceinline__ void blk_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__global__ void __launch_bounds__(kBlkNodesB) k_blk_nodes(const uint32_t *conn, const uint16_t *meta, size_t ne,
                                                          uint32_t eb, const uint32_t *epos, uint32_t *ulist,
                                                          uint32_t *ucount, uint32_t nblk, uint16_t *lconn) {
    constexpr uint32_t CAP = 8 * kBlkNodesE, PER_T = CAP / kBlkNodesB;  // 16 entries per thread
    __shared__ uint32_t v[CAP];  // the sorted vertex ids, then the unique ones in place (16 KB: eight blocks per CU)
    __shared__ uint32_t scan_tmp[kBlkNodesB / 64 + 1];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const size_t e0 = (size_t)b * eb;
    const uint32_t nb = (uint32_t)min((size_t)eb, ne - e0), m = 8 * nb;
    uint32_t P = 1;
    while (P < m) P <<= 1;
    if (b == 0 && tid == 0) ucount[nblk] = 0;  // so that the exclusive scan ends with the total
    for (uint32_t i = tid; i < P; i += kBlkNodesB) v[i] = i < m ? conn[(size_t)(i / nb) * ne + e0 + i % nb] : 0xffffffffu;
    // bitonic sort of v[0..P); compare-exchange c = tid + 256 q touches, for
    // partner distances <= 64, only elements of the 128-element groups
    // 2 (tid / 64) + 8 q and its neighbour, which no other wave touches
    bool prev_local = false;
    for (uint32_t kk = 2; kk <= P; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            const bool local = j <= 64;
            if (local && prev_local)
                blk_wave_sync();
            else
                __syncthreads();
            const uint32_t lj = __ffs(j) - 1;
            for (uint32_t c = tid; c < (P >> 1); c += kBlkNodesB) {
                const uint32_t i = ((c >> lj) << (lj + 1)) | (c & (j - 1)), o = i + j;
                const uint32_t x = v[i], y = v[o];
                if ((i & kk) == 0 ? y < x : x < y) {
                    v[i] = y;
                    v[o] = x;
                }
            }
            prev_local = local;
        }
    }
    __syncthreads();
    // unique, in place: each thread takes its 16 consecutive entries and the
    // one before them into registers; the scan's barriers separate every
    // thread's reads from the compacted writes
    const uint32_t i0 = tid * PER_T;
    uint32_t r[PER_T];
    uint32_t prv = i0 > 0 && i0 - 1 < m ? v[i0 - 1] : 0u, f = 0, fl = 0;
#pragma unroll
    for (int q = 0; q < (int)PER_T; ++q) {
        const uint32_t i = i0 + q;
        r[q] = i < m ? v[i] : 0u;
        const bool head = i < m && (i == 0 || r[q] != prv);
        prv = r[q];
        fl |= head ? 1u << q : 0u;
        f += head ? 1u : 0u;
    }
    uint32_t nu;
    uint32_t w = block_excl_scan_u32(f, scan_tmp, nu);
#pragma unroll
    for (int q = 0; q < (int)PER_T; ++q)
        if ((fl >> q) & 1u) v[w++] = r[q];
    __syncthreads();
    const uint32_t *u = v;
    if (tid == 0) ucount[b] = nu;
    uint32_t *ul = ulist + (size_t)b * 8 * eb;
    for (uint32_t i = tid; i < nu; i += kBlkNodesB) ul[i] = u[i];
    // local index of every vertex, in the child frame of its element
    for (uint32_t i = tid; i < m; i += kBlkNodesB) {
        const uint32_t j = i / nb, k = i % nb;
        const size_t e = e0 + k;
        const uint32_t g = conn[(size_t)j * ne + e];
        uint32_t lo = 0, hi = nu;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (u[mid] < g) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t c = (meta[e] >> 8) & 7u;
        lconn[((size_t)b * eb + epos[e]) * 8 + (j ^ c)] = (uint16_t)(lo * 8u);
    }
}
// Scatters the unique node lists to their block offsets: bnode, pair_block,
// and the number of blocks touching each node.
__global__ void __launch_bounds__(B) k_blk_scatter_nodes(const uint32_t *ulist, uint32_t eb, const uint32_t *boff,
                                                         uint32_t *bnode, uint32_t *pair_block, uint32_t *ncount) {
    const uint32_t b = blockIdx.x, s0 = boff[b], n = boff[b + 1] - s0;
    const uint32_t *ul = ulist + (size_t)b * 8 * eb;
    for (uint32_t i = threadIdx.x; i < n; i += B) {
        const uint32_t g = ul[i];
        bnode[s0 + i] = g;
        pair_block[s0 + i] = b;
        atomicAdd(&ncount[g], 1u);
    }
}
__global__ void k_blk_cmeta(const uint16_t *meta, size_t ne, uint32_t eb, const uint32_t *epos, uint16_t *cmeta) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    uint32_t mt = meta[e], mask = mt & 255u, c = (mt >> 8) & 7u, cm = 0;
    for (uint32_t j = 0; j < 8; ++j)
        if ((mask >> (j ^ c)) & 1u) cm |= 1u << j;
    cmeta[(e / eb) * eb + epos[e]] = (uint16_t)(cm | ((mt >> 11) << 8));
}
// Shared nodes (touched by more than one block) and their node-major partial
// slots, one per block touching the node, blocks in increasing order.
// sh[g]: node g is shared; cnt2[g]: its slot count (0 if not shared).
__global__ void k_shared_count(const uint32_t *ncount, uint32_t nn, uint32_t *sh, uint32_t *cnt2) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    const uint32_t c = ncount[g];
    sh[g] = c > 1 ? 1u : 0u;
    cnt2[g] = c > 1 ? c : 0u;
}
// sidx / base: exclusive scans of sh / cnt2 over the nodes
__global__ void k_shared_index(const uint32_t *ncount, const uint32_t *sidx, const uint32_t *base, uint32_t nn,
                               uint32_t *snode, uint32_t *soff) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn || ncount[g] < 2) return;
    snode[sidx[g]] = g;
    soff[sidx[g]] = base[g];
}
// every shared pair claims a slot of its node (in any order) ...
__global__ void k_shared_claim(const uint32_t *bnode, const uint32_t *pair_block, const uint32_t *ncount,
                               const uint32_t *base, uint64_t np, uint32_t *cursor, uint32_t *tblk,
                               uint32_t *tpair) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t g = bnode[p];
    if (ncount[g] < 2) return;
    const uint32_t q = base[g] + atomicAdd(&cursor[g], 1u);
    tblk[q] = pair_block[p];
    tpair[q] = (uint32_t)p;
}
// ... and each shared node orders its slots by block (few per node)
__global__ void k_shared_order(const uint32_t *soff, uint32_t nsn, uint32_t *tblk, uint32_t *tpair, uint32_t *bdst,
                               uint32_t *bnode) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsn) return;
    const uint32_t a = soff[i], z = soff[i + 1];
    for (uint32_t q = a + 1; q < z; ++q) {
        const uint32_t kb = tblk[q], kp = tpair[q];
        uint32_t r = q;
        while (r > a && tblk[r - 1] > kb) {
            tblk[r] = tblk[r - 1];
            tpair[r] = tpair[r - 1];
            --r;
        }
        tblk[r] = kb;
        tpair[r] = kp;
    }
    for (uint32_t q = a; q < z; ++q) {
        bdst[tpair[q]] = q;
        bnode[tpair[q]] |= 0x80000000u;
    }
}
// Bit 30 of a block node / shared node entry marks a Dirichlet boundary node.
__global__ void k_blk_mark_bdy(uint32_t *ids, uint64_t n, const uint8_t *bdy) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t v = ids[p];
    if (bdy[v & 0x3fffffffu]) ids[p] = v | 0x40000000u;
}
// Node numbering inside a block against LDS bank conflicts.  The matvec adds
// element results into the block's FP64 LDS slots with ds_add_f64, which
// serves 16 lanes per pass and runs without conflicts (8 clocks per wave
// instruction on gfx950) only when their slots differ mod 16; an unordered
// numbering costs about 3x that.  The 16 lanes of one pass are 16 consecutive
// element slots of the block at one record position p, a "lane group".  The
// block's nodes are coloured greedily with 16 colours (colour = slot mod 16):
// each node takes the colour used least so far by the lane groups it appears
// in (ties rotate with the node index), colour c holding
// exactly n/16 + (c < n mod 16) nodes, so that slot = 16 x (rank inside its
// colour) + colour is again a permutation of 0..n-1.  The records and the
// node list are then renumbered.  Deterministic (the use counts are updated
// with atomics); one workgroup per block.  Blocks past this kernel's LDS tables (more than kColMaxN nodes or
// kColMaxE elements) keep the sorted numbering.
// Two waves per block colour 8 nodes per step.  Measured on the lognormal-4M
// DA (matvec back to back / multigrid set-up): one wave, 4 nodes per step,
// 29.5 us / 30.0 ms; two waves 29.7 us / 29.0 ms; four waves 30.0 us / 28.3 ms.
// The LDS tables are sized by the template: <kColSmallN, kColSmallE> (17 KB,
// nine workgroups per CU) takes the blocks that fit it, and <kColMaxN,
// kColMaxE> (33 KB, four per CU) the rest (big = true skips the small ones).
// Both number a block identically.  The colour use counts reuse the degree
// table, and the node-list copy at the end the adjacency table, each once the
// earlier one is dead.
constexpr uint32_t kColMaxN = 2048, kColMaxE = 1024, kColSmallN = 1024, kColSmallE = 512, kColB = 128;
template <uint32_t MAXN, uint32_t MAXE, bool BIG>
__global__ void __launch_bounds__(kColB) k_blk_color(uint16_t *lconn, const uint32_t *boff, uint32_t *bnode,
                                                     size_t ne, uint32_t eb) {
    constexpr uint32_t kCntWords = (MAXE / 16) * 8 * 16 / 4;
    // degree, then fill position; then [lane group][colour] uses so far, one
    // byte each (at most 128 uses per lane group), four to a word for the
    // atomic updates
    __shared__ uint32_t degcnt[MAXN > kCntWords ? MAXN : kCntWords];
    uint32_t *deg = degcnt, *cnt32 = degcnt;
    uint8_t *cnt = reinterpret_cast<uint8_t *>(degcnt);
    __shared__ uint16_t ptr[MAXN + 1];  // adjacency offsets
    // lane groups of each node's uses; then the copy of the node list
    __shared__ uint32_t adjw[4 * MAXE > MAXN ? 4 * MAXE : MAXN];
    uint16_t *adj = reinterpret_cast<uint16_t *>(adjw);
    uint32_t *tmp = adjw;
    __shared__ uint16_t newidx[MAXN];
    __shared__ uint32_t used[16];
    constexpr uint32_t kColG = kColB / 16;
    __shared__ uint32_t pick[kColG];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint32_t s0 = boff[b], n = boff[b + 1] - s0;
    const size_t e0 = (size_t)b * eb;
    if (n > MAXN || eb > MAXE || e0 >= ne) return;  // uniform over the workgroup
    if (BIG && n <= kColSmallN && eb <= kColSmallE) return;  // coloured by the small variant
    const uint32_t nbe = (uint32_t)min((size_t)eb, ne - e0), nuse = 8 * nbe;
    const uint32_t ngroups = ((nbe + 15) / 16) * 8;
    uint16_t *rec = lconn + e0 * 8;
    for (uint32_t i = tid; i < n; i += kColB) deg[i] = 0;
    if (tid < 16) used[tid] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < nuse; t += kColB) atomicAdd(&deg[rec[t] >> 3], 1u);
    __syncthreads();
    if (tid < 64) {  // exclusive scan of the degrees, one chunk per lane of wave 0
        const uint32_t lane = tid;
        const uint32_t ch = (n + 63) / 64, i0 = min(n, lane * ch), i1 = min(n, i0 + ch);
        uint32_t sum = 0;
        for (uint32_t i = i0; i < i1; ++i) sum += deg[i];
        uint32_t incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if ((int)lane >= o) incl += v;
        }
        uint32_t run = incl - sum;
        for (uint32_t i = i0; i < i1; ++i) {
            const uint32_t d = deg[i];
            ptr[i] = (uint16_t)run;
            deg[i] = run;
            run += d;
        }
        if (lane == 63) ptr[n] = (uint16_t)incl;
    }
    __syncthreads();
    for (uint32_t t = tid; t < nuse; t += kColB) {
        const uint32_t q = atomicAdd(&deg[rec[t] >> 3], 1u);
        adj[q] = (uint16_t)((t >> 7) * 8 + (t & 7));  // element slot t/8, lane group (t/8)/16, position t%8
    }
    __syncthreads();
    for (uint32_t i = tid; i < ngroups * 4; i += kColB) cnt32[i] = 0;  // the degrees are dead
    __syncthreads();
    const uint32_t base = n >> 4, extra = n & 15u;
    // kColG nodes per step, one per 16-lane group, taken from the kColG equal
    // parts of the list (far apart, so they rarely share lane groups); lane j
    // of a group scores candidate colour (j + l) mod 16.  Thread 0 then
    // commits the choices in order, moving a choice whose colour has filled up
    // in this step to the next colour with room.
    const uint32_t g = tid >> 4, j = tid & 15u, part = (n + kColG - 1) / kColG;
    for (uint32_t t = 0; t < part; ++t) {
        const uint32_t l = g * part + t;
        const bool has = l < n;
        const uint32_t a = has ? ptr[l] : 0u, z = has ? ptr[l + 1] : 0u, cj = (j + l) & 15u;
        uint32_t pen = 0;
        for (uint32_t q = a; q < z; q += 4) {  // four independent loads per step
            uint32_t u[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) u[r] = q + r < z ? adj[q + r] : 0xffffu;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (u[r] != 0xffffu) pen += cnt[u[r] * 16u + cj];
        }
        uint32_t key = has && used[cj] < base + (cj < extra ? 1u : 0u) ? ((pen << 4) | j) : 0xffffffffu;
        for (int o = 8; o > 0; o >>= 1) key = min(key, (uint32_t)__shfl_xor(key, o, 64));
        if (j == 0) pick[g] = key == 0xffffffffu ? 0xffffffffu : ((key & 15u) + l) & 15u;
        __syncthreads();
        if (tid == 0) {
            for (uint32_t h = 0; h < kColG; ++h) {
                const uint32_t lh = h * part + t;
                if (lh >= n) continue;
                uint32_t c = pick[h];
                if (c == 0xffffffffu || used[c] >= base + (c < extra ? 1u : 0u)) {
                    for (uint32_t d = 0; d < 16; ++d) {
                        const uint32_t c2 = (lh + d) & 15u;
                        if (used[c2] < base + (c2 < extra ? 1u : 0u)) {
                            c = c2;
                            break;
                        }
                    }
                }
                pick[h] = c;
                newidx[lh] = (uint16_t)(used[c] * 16u + c);
                used[c] += 1;
            }
        }
        __syncthreads();
        if (has)
            for (uint32_t q = a + j; q < z; q += 16) {
                const uint32_t idx = adj[q] * 16u + pick[g];
                atomicAdd(&cnt32[idx >> 2], 1u << (8 * (idx & 3u)));
            }
        __syncthreads();
    }
    // Inside each colour the slots follow the node order (ascending ids), so
    // that neighbouring slots hold nearby nodes and a wave's gathers and
    // stores touch fewer lines.  Thread (part p, colour c) counts its colour
    // in the p-th eighth of the list, then renumbers it.
    {
        static_assert(kColB == 8 * 16, "one thread per (eighth, colour)");
        __shared__ uint32_t ccount[8][16];
        const uint32_t c = tid & 15u, p = tid >> 4;
        const uint32_t ch = (n + 7) / 8, i0 = min(n, p * ch), i1 = min(n, i0 + ch);
        uint32_t k = 0;
        for (uint32_t i = i0; i < i1; ++i) k += (newidx[i] & 15u) == c ? 1u : 0u;
        ccount[p][c] = k;
        __syncthreads();
        uint32_t r = 0;
        for (uint32_t q = 0; q < p; ++q) r += ccount[q][c];
        // the colour bits of newidx[] do not change, so other threads' tests stay valid
        for (uint32_t i = i0; i < i1; ++i)
            if ((newidx[i] & 15u) == c) newidx[i] = (uint16_t)((r++) * 16u + c);
        __syncthreads();
    }
    for (uint32_t t = tid; t < nuse; t += kColB) rec[t] = (uint16_t)(newidx[rec[t] >> 3] << 3);
    for (uint32_t i = tid; i < n; i += kColB) tmp[i] = bnode[s0 + i];
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kColB) bnode[s0 + newidx[i]] = tmp[i];
}
// 12-byte element records: the 8 local indices (10 bits each, from lconn's
// byte offsets of 8-byte slots) and cmeta in the top 16 bits.
__global__ void k_blk_pack(const uint16_t *lconn, const uint16_t *cmeta, size_t n, uint32_t *rec) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint32_t v[8];
    for (int j = 0; j < 8; ++j) v[j] = (lconn[8 * e + j] >> 3) & 1023u;
    rec[3 * e] = v[0] | v[1] << 10 | v[2] << 20 | v[3] << 30;
    rec[3 * e + 1] = v[3] >> 2 | v[4] << 8 | v[5] << 18 | v[6] << 28;
    rec[3 * e + 2] = v[6] >> 4 | v[7] << 6 | (uint32_t)cmeta[e] << 16;
}
__global__ void k_blk_maxnodes(const uint32_t *boff, uint32_t nblk, uint32_t *mx) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v = b < nblk ? boff[b + 1] - boff[b] : 0u;
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    if (__lane_id() == 0) atomicMax(mx, v);
}

}  // namespace

void da_build_blocks(DaDevice &da, hipStream_t s, bool colour) {
    const size_t E = da.n_elem, T = 8 * E;
    const uint32_t eb = da.elem_block;
    if (eb == 0 || eb > 8192) {
        dgx_set_error("DA: element block of %u elements is out of range", eb);
        throw DgxError();
    }
    const uint32_t nblk = (uint32_t)((E + eb - 1) / eb);
    uint64_t np;
    DevBuf<uint32_t> pair_block, ncount(da.n_nodes, s);
    HIP_OK(hipMemsetAsync(ncount.p, 0, (size_t)da.n_nodes * 4, s));
    DevBuf<uint32_t> epos(E, s);
    k_blk_order<<<nblk, B, 0, s>>>(da.meta.p, E, eb, epos.p);
    da.lconn.alloc((size_t)nblk * 8 * eb, s);
    da.bnode_off.alloc(nblk + 1, s);
    if (eb <= kBlkNodesE) {
        // per block in LDS: sorted unique nodes and local indices
        DevBuf<uint32_t> ulist((size_t)nblk * 8 * eb, s), ucount(nblk + 1, s);
        k_blk_nodes<<<nblk, kBlkNodesB, 0, s>>>(da.conn.p, da.meta.p, E, eb, epos.p, ulist.p, ucount.p, nblk,
                                                da.lconn.p);
        HIP_OK(hipGetLastError());
        np = scan_exclusive_u32(ucount.p, da.bnode_off.p, nblk + 1, s);
        da.bnode.alloc(np, s);
        pair_block.alloc(np, s);
        k_blk_scatter_nodes<<<nblk, B, 0, s>>>(ulist.p, eb, da.bnode_off.p, da.bnode.p, pair_block.p, ncount.p);
        HIP_OK(hipGetLastError());
    } else {
        // larger blocks: one sort of all (block, node) pairs
        DevBuf<Key> pk(T, s);
        k_blk_pairs<<<grid_for(T, B), B, 0, s>>>(da.conn.p, E, eb, pk.p);
        radix_sort_pairs(pk.p, nullptr, T, s, da.n_nodes, nblk);  // (block, node)
        DevBuf<uint32_t> head(T, s), pos(T, s), bstart(nblk, s);
        k_blk_head<<<grid_for(T, B), B, 0, s>>>(pk.p, T, head.p, bstart.p);
        np = scan_exclusive_u32(head.p, pos.p, T, s);
        da.bnode.alloc(np, s);
        pair_block.alloc(np, s);
        k_blk_compact<<<grid_for(T, B), B, 0, s>>>(pk.p, head.p, pos.p, T, da.bnode.p, pair_block.p, ncount.p);
        k_blk_offsets<<<grid_for(nblk + 1, B), B, 0, s>>>(bstart.p, pos.p, nblk, np, da.bnode_off.p);
        k_blk_lconn<<<grid_for(T, B), B, 0, s>>>(da.conn.p, da.meta.p, E, eb, da.bnode_off.p, da.bnode.p, epos.p,
                                                da.lconn.p);
        HIP_OK(hipGetLastError());
    }
    da.cmeta.alloc((size_t)nblk * eb, s);  // whole blocks, like lconn
    k_blk_cmeta<<<grid_for(E, B), B, 0, s>>>(da.meta.p, E, eb, epos.p, da.cmeta.p);
    if (colour) {
        k_blk_color<kColSmallN, kColSmallE, false><<<nblk, kColB, 0, s>>>(da.lconn.p, da.bnode_off.p, da.bnode.p,
                                                                           E, eb);
        k_blk_color<kColMaxN, kColMaxE, true><<<nblk, kColB, 0, s>>>(da.lconn.p, da.bnode_off.p, da.bnode.p, E, eb);
    }
    HIP_OK(hipGetLastError());
    epos.release();
    // shared nodes and their node-major partial slots
    const uint32_t nn = da.n_nodes;
    uint64_t ns = 0;
    da.bdst.alloc(np, s);
    HIP_OK(hipMemsetAsync(da.bdst.p, 0xff, np * 4, s));
    {
        DevBuf<uint32_t> shf(nn, s), cnt2(nn, s), sidx(nn, s), base(nn, s), tot(2, s);
        k_shared_count<<<grid_for(nn, B), B, 0, s>>>(ncount.p, nn, shf.p, cnt2.p);
        scan_exclusive_u32_async(shf.p, sidx.p, nn, tot.p, s);
        scan_exclusive_u32_async(cnt2.p, base.p, nn, tot.p + 1, s);
        uint32_t ht[2];
        dgx_read_back(ht, tot.p, sizeof ht, s);
        const uint64_t nsn = ht[0];
        ns = ht[1];
        if (ns) {
            da.snode.alloc(nsn, s);
            da.soff.alloc(nsn + 1, s);
            k_shared_index<<<grid_for(nn, B), B, 0, s>>>(ncount.p, sidx.p, base.p, nn, da.snode.p, da.soff.p);
            set_u32_async(da.soff.p + nsn, (uint32_t)ns, s);
            DevBuf<uint32_t> cursor(nn, s), tblk(ns, s), tpair(ns, s);
            HIP_OK(hipMemsetAsync(cursor.p, 0, (size_t)nn * 4, s));
            k_shared_claim<<<grid_for(np, B), B, 0, s>>>(da.bnode.p, pair_block.p, ncount.p, base.p, np, cursor.p,
                                                         tblk.p, tpair.p);
            k_shared_order<<<grid_for(nsn, B), B, 0, s>>>(da.soff.p, (uint32_t)nsn, tblk.p, tpair.p, da.bdst.p,
                                                          da.bnode.p);
            HIP_OK(hipGetLastError());
            da.n_shared = (uint32_t)nsn;
        }
    }
    k_blk_mark_bdy<<<grid_for(np, B), B, 0, s>>>(da.bnode.p, np, da.bdy_node.p);
    if (da.n_shared) k_blk_mark_bdy<<<grid_for(da.n_shared, B), B, 0, s>>>(da.snode.p, da.n_shared, da.bdy_node.p);
    DevBuf<uint32_t> mx(1, s);
    HIP_OK(hipMemsetAsync(mx.p, 0, 4, s));
    k_blk_maxnodes<<<grid_for(nblk, B), B, 0, s>>>(da.bnode_off.p, nblk, mx.p);
    uint32_t hmx = 0;
    dgx_read_back(&hmx, mx.p, 4, s);
    if ((uint64_t)hmx * 8 > 65535) {  // lconn holds 16-bit byte offsets of 8-byte slots
        dgx_set_error("DA: an element block touches %u nodes, more than the 8191 its records can address", hmx);
        throw DgxError();
    }
    if (hmx <= kPipeMaxNodes) {
        const size_t slots = (size_t)nblk * eb;
        da.lrec.alloc(3 * slots, s);
        k_blk_pack<<<grid_for(slots, B), B, 0, s>>>(da.lconn.p, da.cmeta.p, slots, da.lrec.p);
        HIP_OK(hipGetLastError());
    }
    da.n_blocks = nblk;
    da.n_pairs = np;
    da.max_block_nodes = hmx;
    da.partial_slots = ns;
    da.partial.alloc(ns, s);
}
