// dgx_direct.h -- sparse FP64 direct solver of the coarsest multigrid level.
//
// The reference solves its coarsest level with PETSc's sparse LU (bin/
// optionsFBM:57; KSP/PC set-up in src/omg/blockDiag.cpp:780-1047, 1172-1308).
// Here the SPD Dirichlet matrix of the coarsest DA is factorised by a
// multifrontal method on a nested-dissection ordering:
//   - host (symbolic, once per hierarchy): interior-node graph from the
//     element-to-node maps, geometric nested dissection with vertex
//     separators (the cut plane of each part is chosen to minimise the
//     separator), frontal index sets, extend-add maps, and the destination of
//     every matrix entry in the frontal matrices;
//   - device (numeric): element matrices, assembly of the fronts, and per tree
//     level a batched blocked Gauss-Jordan sweep over each front's pivot
//     block, which leaves F11^-1, F11^-1 F12, -F21 F11^-1 and the Schur
//     complement (the update passed to the parent) in place;
//   - device (solve, every V-cycle): a forward sweep up the tree and a
//     backward sweep down it, one launch per level, each a set of dense
//     GEMVs.  Every sum runs in a fixed order, so the solve is deterministic.
// A coarsest level of at most kDenseNodes interior nodes is one front: the
// dense inverse.
#pragma once
#include <memory>
#include <vector>

#include "dgx_da.h"

struct MfFront {
    uint64_t off;       // F: m x m row-major at F + off
    uint64_t tc;        // scratch of the level: T (32 x m) at TC + tc, C (m x 32) after it
    uint32_t m, s;      // front size, pivots (separator nodes)
    uint32_t idx;       // node ids of the front: idx[idx .. idx+m) (S in elimination order, then U)
    uint32_t upd;       // update vector upd[upd .. upd+m-s)
    uint32_t ws;        // forward values of S: ws[ws .. ws+s)
    uint32_t ch0, nch;  // links[ch0 .. ch0+nch): the children
    uint32_t dinv;      // pivot-block inverse scratch index within the level
};
struct MfLink {
    uint32_t child;  // front id
    uint32_t inv;    // inv[inv .. inv + m_parent): child update index of each parent position, or -1
    uint32_t fmap;   // fmap[fmap .. fmap + m_child - s_child): parent position of each child update index
};
struct MfItem {
    uint32_t front, r0;
};

// Host plan (symbolic analysis).  Pure host code: it takes host copies of the
// coarsest DA's element view and node coordinates.
struct MfPlan {
    uint32_t n = 0, ni = 0;                  // nodes, interior nodes
    std::vector<uint32_t> inode;             // interior index -> node id
    std::vector<MfFront> fronts;             // in elimination (post) order
    std::vector<uint32_t> height;            // per front
    std::vector<MfLink> links;
    std::vector<int32_t> inv;
    std::vector<uint32_t> fmap;
    std::vector<uint32_t> idx;               // node ids of all fronts' index sets
    std::vector<uint32_t> a_off, a_src;      // per matrix entry: element-matrix contributions
    std::vector<uint64_t> a_dst;             // per matrix entry: offset in the fronts
    uint64_t front_entries = 0, upd_len = 0;
    uint32_t max_m = 0, max_s = 0;
    // per level (height): fronts, and the solve / gather work items
    std::vector<std::vector<uint32_t>> lvl_fronts;
    std::vector<std::vector<MfItem>> lvl_fwd, lvl_bwd, lvl_gather;
    std::vector<uint32_t> lvl_max_m, lvl_max_s;
    std::vector<uint64_t> lvl_tc;            // scratch doubles of T and C per level
    uint64_t tc_max = 0;
    uint32_t dinv_max = 0;                   // most fronts in one level
};

// Limits of the device kernels: a front's size and separator must fit the
// LDS vectors of the solve sweeps (8192 doubles = 64 KiB).
constexpr uint32_t kMfMaxFront = 8192;
// At most this many interior nodes make one dense front.  Its inverse costs
// more to factorise (about 5 ms at 2800 nodes) and less to apply than the
// multifrontal tree (one GEMV against two launches per tree level and V-cycle:
// lognormal-4M solves 331 -> 366/s, set-up 15.8 -> 21.4 ms), so hierarchies
// solved once (the request server) keep the smaller kDenseNodesOneSolve.
constexpr uint32_t kDenseNodes = 4096;
constexpr uint32_t kDenseNodesOneSolve = 2048;

// Throws DgxError when a front exceeds kMfMaxFront or the fronts exceed
// max_entries doubles.  conn is [8][E] (conn[j*E + e]), meta as DaDevice::meta,
// xyz the node coordinates (3 per node).  Parts of at most `leaf` interior
// nodes are not dissected further; at most `dense` interior nodes in all make
// one front.
MfPlan mf_plan(uint32_t n, size_t E, const uint32_t *conn, const uint16_t *meta, const uint8_t *bdy,
               const uint32_t *xyz, uint64_t max_entries, uint32_t leaf = 128, uint32_t dense = kDenseNodes);

struct CoarseSolver {
    MfPlan plan;  // kept for its sizes and item tables
    uint32_t n = 0, ni = 0, levels = 0, nfronts = 0;
    size_t n_elem = 0;
    DevBuf<MfFront> fronts;
    DevBuf<MfLink> links;
    DevBuf<int32_t> inv;
    DevBuf<uint32_t> fmap, idx, a_off, a_src;
    DevBuf<uint64_t> a_dst;
    std::vector<DevBuf<uint32_t>> lvl_fronts;
    std::vector<DevBuf<MfItem>> lvl_fwd, lvl_bwd, lvl_gather;
    DevBuf<double> F, upd, ws, x;  // fronts, update vectors, forward values, interior solution (by node)
};

// Symbolic analysis of the coarsest DA (host) and upload of the plan.
std::unique_ptr<CoarseSolver> coarse_analyse(const DaDevice &d, hipStream_t s, uint32_t dense = kDenseNodes);
// The same in three steps, so that the host analysis can run beside device
// work: host copies of what the analysis reads (synchronises s), the analysis
// itself (host only: no HIP calls), and the upload of the plan.
struct CoarseInputs {
    uint32_t n = 0;
    size_t E = 0;
    std::vector<uint32_t> conn, xyz;  // conn [8][E]; xyz 3 per node
    std::vector<uint16_t> meta;
    std::vector<uint8_t> bdy;
};
CoarseInputs coarse_download(const DaDevice &d, hipStream_t s);
MfPlan coarse_plan(const CoarseInputs &in, uint32_t dense);
std::unique_ptr<CoarseSolver> coarse_upload(MfPlan &&plan, uint32_t n, size_t E, hipStream_t s);
// Numeric factorisation for the operator lap_fac K + mass_fac M with Dirichlet
// rows; throws DgxError when a pivot is not positive.
void coarse_factor(CoarseSolver &cs, const DaDevice &d, double lap_fac, double mass_fac, hipStream_t s);
// z = A^-1 b (boundary rows: z = b); b and z may be FP32 (inside an FP32
// V-cycle) while the solve runs in FP64.
template <typename T>
void coarse_solve(CoarseSolver &cs, const uint8_t *bdy, const T *b, T *z, hipStream_t s);
