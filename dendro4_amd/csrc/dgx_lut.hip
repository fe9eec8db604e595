// dgx_lut.hip -- compressed element-to-node maps in the reference's form (a
// DA built with compressLut = true).
//
// Encoding (DA::buildNodeList, src/oda/private/odaBuildNlist.cpp:3120-3262),
// per element i with local-buffer node indices nlist[8i..8i+7]:
//   - sort order (m_ucpSortOrders): for an element without hanging vertices,
//     which of the coordinate XORs x ^ (x + sz), y ^ ..., z ^ ... is largest
//     (DA_FLAGS::SortOrderType XYZ 0, YXZ 1, XZY 2, ZXY 3, YZX 4, ZYX 5); for a
//     hanging element its Morton child number;
//   - the 8 indices sorted ascending, nl[0..7]; offsets i - nl[0] and
//     nl[j] - nl[j-1] as unsigned 32-bit values; remainder offset % 256
//     (m_ucpLutRemainders), quotient (unsigned short)(offset >> 8) appended to
//     m_uspLutQuotients when it is not zero, with bit j of m_ucpLutMasks[2i] set.
// Decoding (DA::getNodeIndices, include/oda/oda.tcc:1304-3950): the sorted
// list from remainders and quotients, then unsorted: an element without
// hanging vertices by the fixed pattern of its sort order, a hanging element
// by sorting its 8 vertex points (a hanging vertex j moved to the parent's
// corner j) along the curve.  The quotients are consumed in element order; on
// the device each element finds its first quotient by a scan of the
// per-element counts.
#include "dgx_da.h"
#include "dgx_octree.h"

namespace {

constexpr int LB = 256;

#define LOAD_TABLES(tab)                                                 \
    __shared__ uint8_t s_rc[192], s_mc[192];                             \
    for (int i_ = threadIdx.x; i_ < 192; i_ += blockDim.x) {             \
        s_rc[i_] = (tab).rc[i_];                                          \
        s_mc[i_] = (tab).mc[i_];                                          \
    }                                                                    \
    __syncthreads();

__device__ __forceinline__ void sort8(uint32_t *v) {
#pragma unroll
    for (int a = 1; a < 8; ++a)
#pragma unroll
        for (int b = a; b > 0; --b) {
            const uint32_t lo = min(v[b - 1], v[b]), hi = max(v[b - 1], v[b]);
            v[b - 1] = lo;
            v[b] = hi;
        }
}

__device__ __forceinline__ uint32_t sort_order(uint32_t x, uint32_t y, uint32_t z, uint32_t sz, uint32_t lp,
                                               bool hanging) {
    if (hanging) return 4 * ((z % lp) / sz) + 2 * ((y % lp) / sz) + (x % lp) / sz;
    const uint32_t _x = x ^ (x + sz), _y = y ^ (y + sz), _z = z ^ (z + sz);
    if (_x > _y) return _y > _z ? 5u : (_x > _z ? 4u : 1u);
    return _x > _z ? 3u : (_y > _z ? 2u : 0u);
}

// Per element: remainders, quotient mask, sort order and the quotient count;
// with quot set, the quotients themselves at qoff[i].
__global__ void __launch_bounds__(LB) k_lut_encode(const uint32_t *__restrict__ nlist, const uint8_t *__restrict__ hm,
                                                   const uint32_t *__restrict__ X, const uint32_t *__restrict__ Y,
                                                   const uint32_t *__restrict__ Z, const uint8_t *__restrict__ flags,
                                                   uint32_t md, size_t ne, uint8_t *rem, uint8_t *qmask,
                                                   uint8_t *order, uint32_t *qcnt, const uint32_t *qoff,
                                                   uint16_t *quot) {
    const size_t i = (size_t)blockIdx.x * LB + threadIdx.x;
    if (i >= ne) return;
    uint32_t nl[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) nl[j] = nlist[8 * i + j];
    sort8(nl);
    uint32_t mask = 0, k = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t off = j == 0 ? (uint32_t)i - nl[0] : nl[j] - nl[j - 1];
        const uint16_t q = (uint16_t)(off >> 8);
        if (!quot) rem[8 * i + j] = (uint8_t)(off & 255u);
        if (q) {
            mask |= 1u << j;
            if (quot) quot[qoff[i] + k] = q;
            ++k;
        }
    }
    if (quot) return;
    qmask[i] = (uint8_t)mask;
    qcnt[i] = k;
    const uint32_t d = flags[i] & 31u, sz = 1u << (md - d);
    order[i] = (uint8_t)sort_order(X[i], Y[i], Z[i], sz, 2 * sz, hm[i] != 0);
}

__global__ void __launch_bounds__(LB) k_lut_counts(const uint8_t *qmask, size_t ne, uint32_t *qcnt) {
    const size_t i = (size_t)blockIdx.x * LB + threadIdx.x;
    if (i < ne) qcnt[i] = __popc(qmask[i]);
}

__global__ void __launch_bounds__(LB) k_lut_decode(const uint8_t *__restrict__ rem, const uint8_t *__restrict__ qmask,
                                                   const uint8_t *__restrict__ order,
                                                   const uint16_t *__restrict__ quot, const uint32_t *__restrict__ qoff,
                                                   size_t nq, const uint8_t *__restrict__ hm,
                                                   const uint32_t *__restrict__ X, const uint32_t *__restrict__ Y,
                                                   const uint32_t *__restrict__ Z, const uint8_t *__restrict__ flags,
                                                   uint32_t md, size_t ne, SfcPacked tab, uint32_t *nlist,
                                                   uint32_t *bad) {
    LOAD_TABLES(tab)
    const size_t i = (size_t)blockIdx.x * LB + threadIdx.x;
    if (i >= ne) return;
    // the unsort patterns of the sort orders (oda.tcc:3880-3935): vertex of nn[1..6]
    constexpr uint8_t pat[6][6] = {{1, 2, 3, 4, 5, 6}, {2, 1, 3, 4, 6, 5}, {1, 4, 5, 2, 3, 6},
                                   {4, 1, 5, 2, 6, 3}, {2, 4, 6, 1, 3, 5}, {4, 2, 6, 1, 5, 3}};
    uint32_t nn[8];
    uint32_t qc = qoff[i];
    const uint32_t mask = qmask[i];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        uint32_t v = rem[8 * i + j];
        if (mask & (1u << j)) {
            if (qc >= nq) {
                atomicOr(bad, 1u);
                return;
            }
            v += (uint32_t)quot[qc++] << 8;
        }
        nn[j] = j == 0 ? (uint32_t)i - v : v + nn[j - 1];
    }
    uint32_t *out = nlist + 8 * i;
    const uint32_t h = hm[i], c = order[i];
    if (!h) {
        if (c > 5) {
            atomicOr(bad, 1u);
            return;
        }
        out[0] = nn[0];
#pragma unroll
        for (int k = 0; k < 6; ++k) out[pat[c][k]] = nn[k + 1];
        out[7] = nn[7];
        return;
    }
    const uint32_t d = flags[i] & 31u, sz = 1u << (md - d);
    const uint32_t x = X[i], y = Y[i], z = Z[i];
    const uint32_t px = x - (c & 1u) * sz, py = y - ((c >> 1) & 1u) * sz, pz = z - ((c >> 2) & 1u) * sz;
    Key key[8];
    uint32_t idx[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool hj = (h >> j) & 1u;
        const uint32_t s = hj ? 2 * sz : sz;
        key[j] = dev_encode(s_rc, (hj ? px : x) + (j & 1) * s, (hj ? py : y) + ((j >> 1) & 1) * s,
                            (hj ? pz : z) + ((j >> 2) & 1) * s, md, md);
        idx[j] = j;
    }
    for (int a = 1; a < 8; ++a)
        for (int b = a; b > 0 && key_lt(key[idx[b]], key[idx[b - 1]]); --b) {
            const uint32_t t = idx[b];
            idx[b] = idx[b - 1];
            idx[b - 1] = t;
        }
#pragma unroll
    for (int k = 0; k < 8; ++k) out[idx[k]] = nn[k];
}

}  // namespace

size_t da_lut_compress(const DaDevice &da, LutMaps &m, hipStream_t s) {
    const size_t ne = da.n_elem;
    m.rem.alloc(8 * ne, s);
    m.qmask.alloc(ne, s);
    m.order.alloc(ne, s);
    m.qoff.alloc(ne + 1, s);
    DevBuf<uint32_t> qcnt(ne + 1, s);
    if (ne) {
        k_lut_encode<<<grid_for(ne, LB), LB, 0, s>>>(da.nlist.p, da.mask.p, da.x.p, da.y.p, da.z.p, da.flags.p, da.md,
                                                     ne, m.rem.p, m.qmask.p, m.order.p, qcnt.p, nullptr, nullptr);
        HIP_OK(hipGetLastError());
    }
    m.nq = ne ? scan_exclusive_u32(qcnt.p, m.qoff.p, ne, s) : 0;
    m.quot.alloc(std::max<size_t>(m.nq, 1), s);
    if (ne) {
        k_lut_encode<<<grid_for(ne, LB), LB, 0, s>>>(da.nlist.p, da.mask.p, da.x.p, da.y.p, da.z.p, da.flags.p, da.md,
                                                     ne, nullptr, nullptr, nullptr, nullptr, m.qoff.p, m.quot.p);
        HIP_OK(hipGetLastError());
    }
    return m.nq;
}

void da_lut_decode(const DaDevice &da, LutMaps &m, uint32_t *nlist, hipStream_t s) {
    const size_t ne = da.n_elem;
    if (!ne) return;
    DevBuf<uint32_t> qcnt(ne, s), bad(1, s);
    k_lut_counts<<<grid_for(ne, LB), LB, 0, s>>>(m.qmask.p, ne, qcnt.p);
    m.qoff.alloc(ne + 1, s);
    const uint64_t need = scan_exclusive_u32(qcnt.p, m.qoff.p, ne, s);
    if (need != m.nq) {
        dgx_set_error("compressed maps: the masks call for %llu quotients, %zu given", (unsigned long long)need, m.nq);
        throw DgxError();
    }
    HIP_OK(hipMemsetAsync(bad.p, 0, 4, s));
    k_lut_decode<<<grid_for(ne, LB), LB, 0, s>>>(m.rem.p, m.qmask.p, m.order.p, m.quot.p, m.qoff.p, m.nq, da.mask.p,
                                                 da.x.p, da.y.p, da.z.p, da.flags.p, da.md, ne,
                                                 make_packed(da.curve), nlist, bad.p);
    HIP_OK(hipGetLastError());
    uint32_t hb = 0;
    dgx_read_back(&hb, bad.p, 4, s);
    if (hb) {
        dgx_set_error("compressed maps: a sort order or quotient is out of range");
        throw DgxError();
    }
}
