// dgx_mg.h -- geometric multigrid hierarchy and the CG/V-cycle solve (one rank).
#pragma once
#include <vector>

#include "dgx_da.h"
#include "dgx_direct.h"

// ot::coarsenOctree (src/coarsen.cpp:53-104): every complete set of 8 sibling
// leaves is replaced by its parent, once.
DevBuf<Key> coarsen_build(const Key *in, size_t n, hipStream_t s);

// Interpolation from a coarse level to the next finer one as CSR over fine
// nodes (at most 8 coarse nodes per row), and its transpose as CSR over
// coarse nodes for the restriction.
struct Transfer {
    size_t nf = 0, nc = 0, nnz = 0;
    DevBuf<uint32_t> poff;    // [nf+1]
    DevBuf<uint32_t> pcol;    // [nnz]
    DevBuf<double> pval;      // [nnz]
    DevBuf<float> pval32;     // [nnz]
    DevBuf<uint32_t> roff;    // [nc+1]
    DevBuf<uint32_t> rcol;    // [nnz] fine node ids, ascending per coarse row
    DevBuf<double> rval;      // [nnz]
    DevBuf<float> rval32;     // [nnz]
};

struct MgLevel {
    DevBuf<Key> keys;  // balanced octree of this level (md of the finest)
    size_t n_oct = 0;
    std::unique_ptr<DaDevice> da;
    DevBuf<double> dinv;   // 1 / diag of the Dirichlet operator
    DevBuf<float> dinv32;
    Transfer to_finer;     // interpolation from this level to the finer one (unused on level 0)
};

struct MgWork;  // V-cycle and CG work vectors (dgx_mg.hip)
struct MgHierarchy {
    uint32_t md = 30;
    int curve = 0;
    std::vector<MgLevel> lev;  // lev[0] = finest
    double lap_fac = 1.0, mass_fac = 1.0;
    // coarsest-level sparse FP64 direct solver of the Dirichlet operator
    std::unique_ptr<CoarseSolver> coarse;
    MgWork *work = nullptr;
    ~MgHierarchy();
};

// The caller has waited for every solve on this hierarchy: later solves on
// other streams need not wait for the previous stream.
void mg_mark_idle(MgHierarchy &mg);
// Changes the operator factors: recomputes the level diagonals and the
// coarse inverse, drops captured graphs.
void mg_set_operator(MgHierarchy &mg, double lap_fac, double mass_fac, hipStream_t s);
// A coarsest level of more nodes than this makes the hierarchy coarsen past
// the requested number of levels: the direct solver's analysis, factorisation
// and storage grow faster than the level (fronts of ~2000 nodes and ~1 GB of
// frontal matrices at 75,000 nodes, against ~900 and 0.35 GB at 26,000).
constexpr uint32_t kMaxDirectNodes = 40000;
// With overlap, each level's DA, its diagonal and the coarse solver are built
// on a second host thread and stream while the next levels are coarsened
// (faster when nothing else drives the device; a caller that already runs
// several set-ups at once passes false).  colour as for da_build.
// dense_nodes: the coarsest level's dense-front limit (kDenseNodes, or
// kDenseNodesOneSolve for a hierarchy solved once).
std::unique_ptr<MgHierarchy> mg_build(const Key *finest, size_t n, uint32_t md, int curve, int nlevels,
                                      double lap_fac, double mass_fac, hipStream_t s, bool overlap = true,
                                      bool colour = true, uint32_t dense_nodes = kDenseNodes);

struct SolveOpts {
    int fp32_vcycle = 1;
    int pre = 2, post = 2;
    double omega = 0.8;
    double rtol = 1e-8;
    int maxit = 200;
    int use_graph = 1;  // replay hipGraph captures of the CG step and the V-cycle
};
struct SolveStats {
    int iterations = 0;
    double rel_residual = 0.0;      // ||b - A x|| / ||b|| recomputed in FP64 at the end
    double seconds = 0.0;
};
// Solve A_D u = b on the finest level (device vectors of n_nodes).
void mg_solve(MgHierarchy &mg, const double *b, double *u, const SolveOpts &o, SolveStats *st, hipStream_t s);
// Apply one preconditioner (V-cycle) z = M r, FP64 or FP32 inside.
void mg_vcycle_apply(MgHierarchy &mg, const double *r, double *z, const SolveOpts &o, hipStream_t s);
// FE load vector b = D M f (mass-weighted nodal f, zero on boundary nodes).
void mg_rhs_from_nodal(MgHierarchy &mg, const double *f, double *b, hipStream_t s);
// coarse = R fine / fine = P coarse between levels k and k + 1 (FP64 compact node vectors).
void mg_restrict(MgHierarchy &mg, size_t k, const double *fine, double *coarse, hipStream_t s);
void mg_prolong(MgHierarchy &mg, size_t k, const double *coarse, double *fine, hipStream_t s);
