// dgx_da.h -- device-resident ot::DA (one rank) and its operators.
#pragma once
#include <memory>

#include "dgx_octree.h"

struct DaDevice {
    uint32_t md = 0;  // embedded max depth (D+1)
    int curve = 0;
    size_t n_local = 0, n_elem = 0;
    DevBuf<Key> key;                 // local buffer keys (elements, then pseudo-octants)
    DevBuf<uint32_t> x, y, z;        // anchors in the 2^(D+1) domain
    DevBuf<uint8_t> flags;           // m_ucpOctLevels: level | BOUNDARY(64) | NODE(128)
    DevBuf<uint32_t> nlist;          // m_uiNlist: 8 local-buffer indices per element
    DevBuf<uint8_t> mask;            // hanging mask per element (m_ucpLutMasks[2i+1])
    DevBuf<uint32_t> l2n;            // local index -> compact node id (0xffffffff: not a node)
    uint32_t n_nodes = 0, node_size = 0, bnd_node_size = 0;
    uint32_t n_unresolved = 0, n_ghost_scans = 0;
    // matvec view: conn[j*E + e] compact node id of vertex j; meta = mask | cnum<<8 | level<<11
    DevBuf<uint32_t> conn;
    DevBuf<uint16_t> meta;
    // Block view for the matvec: elements in SFC-contiguous blocks of kElemBlock.
    //   bnode_off[b]..bnode_off[b+1]: the block's unique node ids bnode[] (bit 31: node
    //   also touched by another block), bdst[]: its partial slot when shared.
    //   Inside a block the elements are reordered (no hanging vertex first); slot k
    //   of block b holds lconn[(b*elem_block + k)*8 + j] = 8 x (local index of vertex
    //   j ^ c in the block's list), c the child number -- the "child frame", where
    //   the parent corner is vertex 0 -- and cmeta[b*elem_block + k] = hanging mask
    //   in the child frame | original level << 8.  Both hold whole blocks (the slots past
    //   n_elem in the last block are allocated and never used).
    //   shared nodes: snode[s], partial slots soff[s]..soff[s+1] (ordered by block).
    uint32_t elem_block = 512;  // elements per block (a block may touch at most 8191 nodes)
    uint32_t n_blocks = 0, max_block_nodes = 0, n_shared = 0;
    uint64_t n_pairs = 0, partial_slots = 0;
    DevBuf<uint32_t> bnode_off, bnode, bdst, soff, snode;
    DevBuf<uint16_t> lconn, cmeta;
    // The same records packed into 12 bytes for blocks of at most kPipeMaxNodes
    // nodes (the pipelined matvec): three words per element slot, bits 10j..10j+9
    // the local index of vertex j, bits 80..95 cmeta.  Empty when a block
    // touches more nodes.
    DevBuf<uint32_t> lrec;
    DevBuf<double> partial;  // partial sums of shared nodes (float or double view), one per matvec in flight
    // Per node: local-buffer index of its octant, and 1 if the node lies on the
    // domain boundary (homogeneous Dirichlet rows, omgJac.C:958-1057).
    DevBuf<uint32_t> n2l;
    DevBuf<uint8_t> bdy_node;
    uint32_t n_bdy_nodes = 0;
    ~DaDevice();
};

// Largest block node list the pipelined matvec (and its 10-bit packed records) takes.
constexpr uint32_t kPipeMaxNodes = 1024;

// colour: renumber each block's nodes against LDS bank conflicts (k_blk_color).
// It makes every later matvec faster and costs about 40 matvecs once, so a
// DA used for few matvecs (one served solve) is built without it.
std::unique_ptr<DaDevice> da_build(const Key *oct, size_t n, uint32_t md0, int curve, hipStream_t s,
                                   bool colour = true);
// Builds the matvec block view from conn (called by da_build).
void da_build_blocks(DaDevice &da, hipStream_t s, bool colour = true);

// y = A x with A = lapFac * stiffness + massFac * mass (type-2 hanging stencils),
// FP64 on compact node vectors (Jacobian1MatMult, odaJac.C:275-366).
void da_matvec_f64(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac, hipStream_t s);
void da_matvec_f32(const DaDevice &da, const float *x, float *y, float lap_fac, float mass_fac, hipStream_t s);
// The FP64 matvec with the shared nodes' partial sums in a static device array
// of kNoAllocScratchBytes (one per device) instead of the DA's buffer: no
// allocation.  Throws DgxError when the DA's partial slots exceed the array.
// Calls that overlap on one device race on the array.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
void da_matvec_f64_noalloc(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac,
                           hipStream_t s);
// Takes the array's device address for the current device (once per device).
void da_noalloc_init();
// Diagonal of the same operator (Jacobian1MatGetDiagonal, odaJac.C:160-227).
void da_diagonal_f64(const DaDevice &da, double *diag, double lap_fac, double mass_fac, hipStream_t s);
// Homogeneous-Dirichlet operator (DirichletJacobianMatMult, omgJac.C:958-1057):
// boundary rows are the identity and boundary columns are dropped.
void da_matvec_dirichlet_f64(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac,
                             hipStream_t s);
void da_matvec_dirichlet_f32(const DaDevice &da, const float *x, float *y, float lap_fac, float mass_fac,
                             hipStream_t s);
// Smoother and residual of the Dirichlet operator with the update fused into
// the matvec's stores: x += omega dinv (b - A x) in place (scratch: n_nodes
// entries, used only for DAs whose blocks exceed 1024 nodes), r = b - A x.
template <typename T>
void da_dirichlet_jacobi(const DaDevice &da, T *x, const T *b, const T *dinv, T omega, T lap_fac, T mass_fac,
                         T *scratch, hipStream_t s);
template <typename T>
void da_dirichlet_residual(const DaDevice &da, const T *x, const T *b, T *r, T lap_fac, T mass_fac, hipStream_t s);
// Compressed element-to-node maps in the reference's form (compressLut,
// odaBuildNlist.cpp:3120-3262): 8 remainders, a quotient mask and a sort
// order per element, the quotients in element order, and (derived here) each
// element's first quotient.
struct LutMaps {
    DevBuf<uint8_t> rem, qmask, order;
    DevBuf<uint32_t> qoff;
    DevBuf<uint16_t> quot;
    size_t nq = 0;
};
// Encodes da's nlist; returns the number of quotients.
size_t da_lut_compress(const DaDevice &da, LutMaps &m, hipStream_t s);
// DA::getNodeIndices over compressed maps (oda.tcc:1304-3950) for every
// element: nlist (device, 8 per element) receives the decoded indices.
void da_lut_decode(const DaDevice &da, LutMaps &m, uint32_t *nlist, hipStream_t s);
// Element matrices of lap_fac K + mass_fac M in each element's child frame:
// vals[e*64 + k*8 + j] is the contribution to A[id[j^c], id[k^c]], id the
// element's 8 node ids (conn) and c its child number (meta).
void da_element_matrices(const DaDevice &da, double lap_fac, double mass_fac, double *vals, hipStream_t s);
