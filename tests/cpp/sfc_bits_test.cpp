// Host check of the key arithmetic shared by the octree kernels
// (dendro4_amd/csrc/dgx_common.h, dgx_octree.h): the bit-spread Morton
// encode/decode against the per-level definition (digit t of a key = bit
// md - t of z, y, x), and the two-levels-per-read conversion of Morton keys
// to curve keys (morton_to_curve_rot) against the per-level table walk
// sfc_encode, for both curves, every maxDepth 1..31 and every level.
#include <cstdio>
#include <random>

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
    printf("%ld cases, %ld failures\n", n, bad);
    return bad ? 1 : 0;
}
