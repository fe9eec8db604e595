// dgx_octree.h -- octree construction and balance (device side, internal API).
#pragma once
#include "dgx_common.h"

// Packed per-curve tables staged into LDS by the kernels that encode/decode:
//   rc[r*8+m] = rank(r,m) | crot(r,m) << 3        (encode: Morton child -> rank)
//   mc[r*8+k] = morton(r,k) | crot(r,morton) << 3 (decode: rank -> Morton child)
struct SfcPacked {
    uint8_t rc[192];
    uint8_t mc[192];
};
SfcPacked make_packed(int curve);

__device__ inline Key dev_encode(const uint8_t *rc, uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    Key k{0ull, 0ull};
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        uint32_t b = md - t;
        uint32_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        uint32_t v = rc[rot * 8 + m];
        rot = v >> 3;
        if (t <= 21) k.hi |= (uint64_t)(v & 7u) << (63 - 3 * t);
        else k.lo |= (uint64_t)(v & 7u) << (63 - 3 * (t - 21));
    }
    k.lo |= lev;
    return k;
}
// dev_encode that also returns the curve's rotation state below the octant,
// so that its children's keys need one table read each (child_key).
__device__ inline Key dev_encode_rot(const uint8_t *rc, uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md,
                                     uint32_t &rot) {
    Key k{0ull, 0ull};
    rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        uint32_t b = md - t;
        uint32_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        uint32_t v = rc[rot * 8 + m];
        rot = v >> 3;
        if (t <= 21) k.hi |= (uint64_t)(v & 7u) << (63 - 3 * t);
        else k.lo |= (uint64_t)(v & 7u) << (63 - 3 * (t - 21));
    }
    k.lo |= lev;
    return k;
}
// Two curve levels per table read: rc2[r*64 + (m1 << 3 | m2)] = rank1 << 3 |
// rank2 | rot2 << 6 for Morton children m1 (level t) and m2 (level t+1) below
// rotation state r.  Filled in LDS from rc by the workgroup.
__host__ __device__ inline uint16_t rc2_entry(const uint8_t *rc, uint32_t i) {
    const uint32_t r = i >> 6, m1 = (i >> 3) & 7u, m2 = i & 7u;
    const uint32_t v1 = rc[r * 8 + m1], v2 = rc[(v1 >> 3) * 8 + m2];
    return (uint16_t)(((v1 & 7u) << 3) | (v2 & 7u) | ((v2 >> 3) << 6));
}
// The output-curve key of the octant with Morton key mk, and the curve's
// rotation state below it (as dev_encode_rot of its anchor).  A Morton
// digit pair of levels t, t+1 is one 6-bit field, and the curve's digit pair
// goes to the same field.
__host__ __device__ inline Key morton_to_curve_rot(const uint16_t *rc2, const uint8_t *rc, const Key &mk,
                                                   uint32_t &rot) {
    const uint32_t lev = key_level(mk);
    Key k{0ull, 0ull};
    uint32_t r = 0, t = 1;
    for (; t + 1 <= lev && t + 1 <= 20; t += 2) {  // pairs (1,2) .. (19,20) in the hi word
        const uint32_t sh = 60 - 3 * t, v = rc2[r * 64 + ((uint32_t)(mk.hi >> sh) & 63u)];
        r = v >> 6;
        k.hi |= (uint64_t)(v & 63u) << sh;
    }
    for (; t <= lev && t <= 21; ++t) {  // an odd last level, or level 21
        const uint32_t v = rc[r * 8 + key_digit(mk, t)];
        r = v >> 3;
        k.hi |= (uint64_t)(v & 7u) << (63 - 3 * t);
    }
    for (; t + 1 <= lev; t += 2) {  // pairs (22,23) .. in the lo word
        const uint32_t sh = 60 - 3 * (t - 21), v = rc2[r * 64 + ((uint32_t)(mk.lo >> sh) & 63u)];
        r = v >> 6;
        k.lo |= (uint64_t)(v & 63u) << sh;
    }
    if (t <= lev) {
        const uint32_t v = rc[r * 8 + key_digit(mk, t)];
        r = v >> 3;
        k.lo |= (uint64_t)(v & 7u) << (63 - 3 * (t - 21));
    }
    k.lo |= lev;
    rot = r;
    return k;
}
// Curve keys of points near a reference octant, from the reference's path.
// path_walk encodes the octant at level d anchored at (x, y, z) and records
// the rotation state before every level t = 1..d+1 in rots[t * stride].
__host__ __device__ inline Key path_walk(const uint8_t *rc, uint32_t x, uint32_t y, uint32_t z, uint32_t d,
                                         uint32_t md, uint8_t *rots, uint32_t stride) {
    Key k{0ull, 0ull};
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= d; ++t) {
        rots[t * stride] = (uint8_t)rot;
        const uint32_t b = md - t;
        const uint32_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        const uint32_t v = rc[rot * 8 + m];
        rot = v >> 3;
        key_set_digit(k, t, v & 7u);
    }
    rots[(d + 1) * stride] = (uint8_t)rot;
    k.lo |= d;
    return k;
}
// Digits of levels s..md of the curve path whose Morton digits are all 0 (a
// point of the level-(s-1) grid below level s-1), starting in rotation state r
// before level s; the level field is 0.  Levels past md give no digits.
__host__ __device__ inline Key zero_tail(const uint8_t *rc, uint32_t r, uint32_t s, uint32_t md) {
    Key k{0ull, 0ull};
    for (uint32_t t = s; t <= md; ++t) {
        const uint32_t v = rc[r * 8];
        r = v >> 3;
        key_set_digit(k, t, v & 7u);
    }
    return k;
}
constexpr uint32_t kTailCols = 33;  // start levels 0..32 of the zero_tail table [24][kTailCols]
// The key at level lt of the point (px, py, pz) of the level-g grid
// (coordinates multiples of 2^(md-g), taken mod 2^md), g <= d, from the path
// of the level-d octant anchored at (x, y, z): its key kref and rotation
// states rots (path_walk).  Levels above the first one at which the point and
// the anchor differ are the reference's digits; the levels from there down
// to g are walked; the levels g+1..lt, where the point's Morton digits are
// 0, come from tail[rot * kTailCols + g + 1] = zero_tail(rc, rot, g + 1, md)
// (needed only when lt > g, and then lt must be md).  Equals the per-level
// encode of the point at level lt.
__host__ __device__ inline Key key_from_path(const uint8_t *rc, const Key &kref, const uint8_t *rots, uint32_t stride,
                                             uint32_t x, uint32_t y, uint32_t z, uint32_t px, uint32_t py,
                                             uint32_t pz, uint32_t g, uint32_t lt, uint32_t md, const Key *tail) {
    const uint32_t dom = md >= 32 ? 0xffffffffu : (1u << md) - 1u;
    const uint32_t diff = ((x ^ px) | (y ^ py) | (z ^ pz)) & dom;
    uint32_t p = g;  // levels 1..p are the reference's digits
    if (diff) {
        const uint32_t first = md - (31u - (uint32_t)__builtin_clz(diff));  // first level that differs
        if (first - 1 < p) p = first - 1;
    }
    Key k = key_truncate(kref, p);
    k.lo &= ~31ull;
    uint32_t rot = rots[(p + 1) * stride];
    for (uint32_t t = p + 1; t <= g; ++t) {
        const uint32_t b = md - t;
        const uint32_t m = (((pz >> b) & 1u) << 2) | (((py >> b) & 1u) << 1) | ((px >> b) & 1u);
        const uint32_t v = rc[rot * 8 + m];
        rot = v >> 3;
        key_set_digit(k, t, v & 7u);
    }
    if (lt > g) {
        const Key &tl = tail[rot * kTailCols + g + 1];
        k.hi |= tl.hi;
        k.lo |= tl.lo;
    }
    k.lo |= lt;
    return k;
}
// Key of Morton child m (bit 0: +x, 1: +y, 2: +z) of the octant with key p
// and rotation state rot (from dev_encode_rot); equals dev_encode of the child.
__device__ inline Key child_key(const uint8_t *rc, Key p, uint32_t rot, uint32_t m) {
    const uint32_t t = key_level(p) + 1;
    key_set_digit(p, t, rc[rot * 8 + m] & 7u);
    p.lo = (p.lo & ~31ull) | t;
    return p;
}
__device__ inline void dev_decode(const uint8_t *mc, const Key &k, uint32_t md, uint32_t &x, uint32_t &y, uint32_t &z,
                                  uint32_t &lev) {
    lev = key_level(k);
    x = y = z = 0;
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        uint32_t v = mc[rot * 8 + key_digit(k, t)];
        uint32_t m = v & 7u, b = md - t;
        x |= (m & 1u) << b;
        y |= ((m >> 1) & 1u) << b;
        z |= ((m >> 2) & 1u) << b;
        rot = v >> 3;
    }
}

struct ReplayResult {
    uint32_t nout, end, err, next_t;  // next_t: first textbook leaf at or inside resync
    Key resync;
};

struct P2OStats {
    uint64_t textbook_leaves = 0, triggers = 0, replays = 0, replayed_leaves = 0, leaves = 0;
};
struct BalStats {
    uint64_t levels = 0, seeds = 0, leaves = 0;
};

// Sorted leaf keys of the octree built from n points (xyz interleaved, device).
DevBuf<Key> p2o_build(const double *d_xyz, size_t n, uint32_t md, uint32_t max_num_pts, int curve, hipStream_t s,
                      P2OStats *stats, const double *glens = nullptr);
// Sorted leaf keys of the minimal 2:1 (face, edge, corner) balanced refinement.
DevBuf<Key> balance_build(const Key *in, size_t n, uint32_t md, int curve, hipStream_t s, BalStats *stats);
// Keys <-> ot::TreeNode arrays (7 x uint32 per octant, device).
void export_treenodes(const Key *k, size_t n, uint32_t md, int curve, uint32_t *d_tn, hipStream_t s);
DevBuf<Key> import_treenodes(const uint32_t *d_tn, size_t n, uint32_t md, int curve, bool sort, hipStream_t s);
