// dgx_tables.cpp -- space-filling-curve tables.
//
// Hilbert: 24 rotations of the 3-D curve.  For rotation r the reference stores
// a 16-character string whose first 8 characters give the Morton index of the
// children in curve order and whose last 8 give the curve rank of each Morton
// child, plus the rotation entered in each Morton child (HILBERT_TABLE); values
// from src/hcurvedata.cpp:329-546.  They are kept here as data in that form and
// expanded into rank/morton/crot arrays at start-up; a CPU test checks that the
// two halves are inverse permutations and that consecutive children (and the
// sub-curves across child boundaries) are face-adjacent, i.e. that the tables
// describe a Hilbert curve.
// Morton: identity ranks and a single rotation (src/hcurvedata.cpp:548-563).
#include <cstring>

#include "dgx_common.h"
#include "dgx_octree.h"

namespace {

const char *kHilbertPerm =
    "0231576403127465" "0451376203741265" "0264573107163425" "6204513725163407"
    "3751046243705261" "5731026443527061" "6237510465127403" "0462375107341625"
    "6402315725341607" "6457310265741203" "0154673201763245" "3201546723105467"
    "6754013245763201" "5764023147563021" "3267540167105423" "5104623721563047"
    "3762045147305621" "6732015445327601" "5401326723541067" "3157640261705243"
    "5137620461527043" "0132675401327645" "5467320167541023" "3102645721305647";

const uint8_t kHilbertChildRot[192] = {
    1,  3,  2,  0,  6,  4,  5,  0,  0,  8,  9,  5,  7,  1,  4,  1,  10, 14, 0,  13, 11, 12, 2,  2,
    3,  3,  6,  15, 14, 10, 12, 11, 18, 17, 12, 11, 4,  4,  1,  16, 11, 12, 5,  5,  17, 18, 0,  13,
    20, 6,  3,  6,  0,  19, 9,  2,  21, 22, 18, 17, 1,  16, 7,  7,  8,  8,  22, 21, 9,  23, 17, 18,
    19, 9,  1,  20, 8,  9,  6,  7,  2,  21, 19, 12, 23, 10, 13, 10, 11, 6,  14, 4,  11, 1,  18, 3,
    15, 12, 4,  10, 16, 12, 3,  17, 3,  16, 13, 2,  4,  15, 13, 5,  22, 2,  11, 19, 14, 9,  14, 0,
    15, 20, 15, 3,  19, 13, 2,  23, 8,  13, 5,  23, 16, 7,  16, 4,  23, 17, 13, 17, 5,  21, 8,  12,
    18, 9,  18, 0,  22, 5,  11, 8,  9,  23, 10, 14, 19, 19, 22, 21, 6,  15, 20, 20, 21, 22, 14, 10,
    7,  10, 15, 21, 20, 17, 16, 21, 14, 7,  22, 6,  18, 20, 22, 1,  23, 19, 20, 16, 23, 8,  7,  15};

struct TableInit {
    SfcTables t[2];
    TableInit() {
        for (int r = 0; r < 24; ++r)
            for (int i = 0; i < 8; ++i) {
                t[0].morton[r * 8 + i] = (uint8_t)(kHilbertPerm[r * 16 + i] - '0');
                t[0].rank[r * 8 + i] = (uint8_t)(kHilbertPerm[r * 16 + 8 + i] - '0');
                t[0].crot[r * 8 + i] = kHilbertChildRot[r * 8 + i];
                t[1].morton[r * 8 + i] = (uint8_t)i;
                t[1].rank[r * 8 + i] = (uint8_t)i;
                t[1].crot[r * 8 + i] = 0;
            }
    }
};
const TableInit &tables() {
    static TableInit ti;
    return ti;
}

}  // namespace

const SfcTables &host_sfc(int curve) { return tables().t[curve == CURVE_MORTON ? 1 : 0]; }

extern "C" int dgx_sfc_tables(int curve, uint8_t *rank192, uint8_t *morton192, uint8_t *crot192) {
    if (curve != CURVE_HILBERT && curve != CURVE_MORTON) return -1;
    const SfcTables &t = host_sfc(curve);
    if (rank192) memcpy(rank192, t.rank, 192);
    if (morton192) memcpy(morton192, t.morton, 192);
    if (crot192) memcpy(crot192, t.crot, 192);
    return 0;
}

SfcPacked make_packed(int curve) {
    const SfcTables &T = host_sfc(curve);
    SfcPacked p;
    for (int r = 0; r < 24; ++r)
        for (int m = 0; m < 8; ++m) {
            p.rc[r * 8 + m] = (uint8_t)(T.rank[r * 8 + m] | (T.crot[r * 8 + m] << 3));
            uint8_t mm = T.morton[r * 8 + m];  // m plays the role of the rank here
            p.mc[r * 8 + m] = (uint8_t)(mm | (T.crot[r * 8 + mm] << 3));
        }
    return p;
}
