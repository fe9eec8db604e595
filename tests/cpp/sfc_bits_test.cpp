// Host check of the key arithmetic shared by the octree kernels
// (dendro4_amd/csrc/dgx_common.h, dgx_octree.h): the bit-spread Morton
// encode/decode against the per-level definition (digit t of a key = bit
// md - t of z, y, x), and the two-levels-per-read conversion of Morton keys
// to curve keys (morton_to_curve_rot) against the per-level table walk
// sfc_encode, for both curves, every maxDepth 1..31 and every level.  Also
// the keys the DA build derives from an element's path (path_walk,
// key_from_path with the zero_tail table): the element's vertices and its
// parent's vertices at level md, its neighbours' parents at the level above,
// points of the coarser grids anywhere in the domain, and vertices one cell
// past the domain (taken mod 2^md), against sfc_encode.
#include <cstdio>
#include <random>
#include <vector>

#include "dgx_octree.h"

static Key morton_by_level(uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    Key k{0ull, 0ull};
    for (uint32_t t = 1; t <= lev; ++t) {
        const uint32_t b = md - t;
        const uint64_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        if (t <= 21) k.hi |= m << (63 - 3 * t);
        else k.lo |= m << (63 - 3 * (t - 21));
    }
    k.lo |= lev;
    return k;
}

int main() {
    std::mt19937_64 rng(7);
    long n = 0, bad = 0;
    for (int curve = 0; curve < 2; ++curve) {
        const SfcTables &T = host_sfc(curve);
        const SfcPacked P = make_packed(curve);
        uint16_t rc2[24 * 64];
        for (uint32_t i = 0; i < 24 * 64; ++i) rc2[i] = rc2_entry(P.rc, i);
        for (uint32_t md = 1; md <= 31; ++md)
            for (uint32_t lev = 0; lev <= md; ++lev)
                for (int r = 0; r < 400; ++r) {
                    // arbitrary words (bits outside md and finer than lev are ignored)
                    uint32_t x = (uint32_t)rng(), y = (uint32_t)rng(), z = (uint32_t)rng();
                    const Key ref = morton_by_level(x, y, z, lev, md);
                    const Key got = morton_encode(x, y, z, lev, md);
                    ++n;
                    if (got.hi != ref.hi || got.lo != ref.lo) {
                        if (bad++ < 8) printf("morton_encode md=%u lev=%u\n", md, lev);
                        continue;
                    }
                    // decode, also with junk in the digits finer than the level
                    Key k = ref;
                    if (r & 1) {
                        if (lev < 21) k.hi |= rng() & ((1ull << (63 - 3 * lev)) - 1ull);
                        k.lo |= rng() & (lev > 21 ? ((1ull << (126 - 3 * lev)) - 1ull) : ~0ull) & ~31ull;
                    }
                    uint32_t dx, dy, dz, dl;
                    morton_decode(k, md, dx, dy, dz, dl);
                    const uint32_t keep = (uint32_t)(((1ull << md) - 1ull) & ~((1ull << (md - lev)) - 1ull));
                    if (dl != lev || dx != (x & keep) || dy != (y & keep) || dz != (z & keep)) {
                        if (bad++ < 8) printf("morton_decode md=%u lev=%u\n", md, lev);
                        continue;
                    }
                    uint32_t rot = 0;
                    const Key c = morton_to_curve_rot(rc2, P.rc, ref, rot);
                    const Key e = sfc_encode(T, x & keep, y & keep, z & keep, lev, md);
                    uint32_t rot_ref = 0;  // rotation below the octant, by the per-level walk
                    for (uint32_t t = 1; t <= lev; ++t) rot_ref = T.crot[rot_ref * 8 + key_digit(ref, t)];
                    if (c.hi != e.hi || c.lo != e.lo || rot != rot_ref) {
                        if (bad++ < 8) printf("morton_to_curve_rot curve=%d md=%u lev=%u\n", curve, md, lev);
                    }
                }
    }
    // keys from an element's path
    for (int curve = 0; curve < 2; ++curve) {
        const SfcTables &T = host_sfc(curve);
        const SfcPacked P = make_packed(curve);
        for (uint32_t md = 2; md <= 31; ++md) {
            std::vector<Key> tail(24 * kTailCols);
            for (uint32_t r = 0; r < 24; ++r)
                for (uint32_t s0 = 0; s0 < kTailCols; ++s0) tail[r * kTailCols + s0] = zero_tail(P.rc, r, s0, md);
            const uint32_t dom = md == 32 ? 0xffffffffu : (1u << md) - 1u;
            for (uint32_t d = 1; d <= md; ++d)
                for (int r = 0; r < 60; ++r) {
                    const uint32_t sh = md - d, sz = 1u << sh;
                    const uint32_t x = ((uint32_t)rng() & dom) >> sh << sh, y = ((uint32_t)rng() & dom) >> sh << sh,
                                   z = ((uint32_t)rng() & dom) >> sh << sh;
                    uint8_t rots[33];
                    const Key kref = path_walk(P.rc, x, y, z, d, md, rots, 1);
                    const Key e0 = sfc_encode(T, x, y, z, d, md);
                    ++n;
                    if (kref.hi != e0.hi || kref.lo != e0.lo) {
                        if (bad++ < 8) printf("path_walk curve=%d md=%u d=%u\n", curve, md, d);
                        continue;
                    }
                    auto check = [&](uint32_t px, uint32_t py, uint32_t pz, uint32_t g, uint32_t lt, const char *what) {
                        const Key got = key_from_path(P.rc, kref, rots, 1, x, y, z, px, py, pz, g, lt, md, tail.data());
                        const Key ref = sfc_encode(T, px & dom, py & dom, pz & dom, lt, md);
                        ++n;
                        if (got.hi != ref.hi || got.lo != ref.lo) {
                            if (bad++ < 8) printf("key_from_path %s curve=%d md=%u d=%u g=%u lt=%u\n", what, curve, md, d, g, lt);
                        }
                    };
                    const uint32_t ps = sh + 1;
                    const uint32_t parX = d >= 1 ? x >> ps << ps : 0, parY = d >= 1 ? y >> ps << ps : 0,
                                   parZ = d >= 1 ? z >> ps << ps : 0;
                    for (uint32_t j = 0; j < 8; ++j) {
                        const uint32_t bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
                        // vertex of the element (k_da_nlist), level md
                        check(x + bx * sz, y + by * sz, z + bz * sz, d, md, "vertex");
                        // vertex of the parent (k_da_nlist), level md
                        check(parX + bx * (sz << 1), parY + by * (sz << 1), parZ + bz * (sz << 1), d - 1, md, "parent vertex");
                    }
                    if (d >= 2) {
                        // parents of the same-level neighbours (k_da_flag_nodes), level d - 1
                        const uint32_t nb[3] = {x >= sz ? x - sz : x, y >= sz ? y - sz : y, z >= sz ? z - sz : z};
                        for (uint32_t j = 1; j < 8; ++j) {
                            const uint32_t a = j & 1 ? nb[0] : x, b = j & 2 ? nb[1] : y, c = j & 4 ? nb[2] : z;
                            check(a >> ps << ps, b >> ps << ps, c >> ps << ps, d - 1, d - 1, "neighbour parent");
                        }
                    }
                    // points of a coarser grid anywhere in the domain, both target levels
                    const uint32_t g = d - (uint32_t)(rng() % (d + 1)), gs = md - g;
                    const uint32_t qx = gs >= 32 ? 0 : ((uint32_t)rng() & dom) >> gs << gs;
                    const uint32_t qy = gs >= 32 ? 0 : ((uint32_t)rng() & dom) >> gs << gs;
                    const uint32_t qz = gs >= 32 ? 0 : ((uint32_t)rng() & dom) >> gs << gs;
                    check(qx, qy, qz, g, g, "grid point");
                    check(qx, qy, qz, g, md, "grid point at md");
                }
        }
    }
    printf("%ld cases, %ld failures\n", n, bad);
    return bad ? 1 : 0;
}
