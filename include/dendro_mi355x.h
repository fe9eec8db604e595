/*
 * dendro_mi355x.h -- C ABI of the MI355X octree-FEM path.
 *
 * Drop-in boundary for the one-rank (MPI_COMM_SELF) hot path of Dendro4:
 * points -> ot::points2Octree -> ot::balanceOctree -> ot::DA -> -lap+I matvec
 * -> DAMG-preconditioned CG.  Every entry point returns 0 on success and a
 * negative value on error; dgx_last_error() then describes the failure.
 * Handles own device-resident state (keys, DA maps, vectors) on the calling
 * process's current device; host arrays are copied only when a call asks for
 * them.  Not thread-safe: calls on one process are serialised by the caller,
 * as in the reference (global Hilbert tables, hcurvedata.h:42-43).
 *
 * TreeNode arrays are 7 x uint32 per octant with the ot::TreeNode layout
 * (include/TreeNode.h:42): x, y, z, level|flags, weight, dim, maxDepth.
 */
#ifndef DENDRO_MI355X_H
#define DENDRO_MI355X_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgx_octree_s *dgx_octree;
typedef struct dgx_da_s *dgx_da;

/* Curves: DGX_CURVE_HILBERT matches a reference build with HILBERT_ORDERING
 * (its default, CMakeLists.txt:48); DGX_CURVE_MORTON matches one without. */
enum { DGX_CURVE_HILBERT = 0, DGX_CURVE_MORTON = 1 };

/* ---- runtime ------------------------------------------------------------ */
/* Selects the device and creates the library stream and memory pool.
 * Replaces the one-time set-up of _InitializeHcurve (hcurvedata.h:45) and
 * DA_Initialize (odaUtils.cpp:3352). */
int dgx_init(int device);
const char *dgx_last_error(void);
/* The hipStream_t all library work is queued on. */
void *dgx_stream(void);
int dgx_synchronize(void);
int dgx_device_malloc(void **p, size_t bytes);
int dgx_device_free(void *p);
int dgx_memcpy_htod(void *dst, const void *src, size_t bytes);
int dgx_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int dgx_memset_device(void *dst, int value, size_t bytes);
/* Device memory (no reference counterpart): out[0] device total, [1] in use on
 * the device (total - free, all allocators), [2] the library's allocations in
 * use, [3] their peak since start or the last reset (reset_peak != 0 resets it
 * after reading), [4] bytes the library's block cache holds. */
int dgx_memory_info(uint64_t out[5], int reset_peak);
/* No reference counterpart: returns the blocks the library's block cache
 * holds to the device pool (except those cached for a stream inside a graph
 * capture), waits for the device, and releases the pool's unused memory to
 * the device. */
int dgx_trim_cache(void);

/* ---- octree ------------------------------------------------------------- */
/* ot::points2Octree on one rank (include/octUtils.h:633-634; construct.cpp:
 * 759-774 -> points2OctreeSeq :911-974 -> p2oLocal :976-1077).  xyz holds n
 * interleaved points in [0,1)^3 (gLens = 1); on_device selects whether xyz is
 * a device pointer.  The result is sorted, linear and complete. */
int dgx_points_to_octree(const double *xyz, size_t n, unsigned max_depth, unsigned max_num_pts, int curve,
                         int on_device, dgx_octree *out);
/* Same with the domain lengths gLens[3] of points2Octree (octUtils.h:633):
 * a point quantises to (unsigned)(x * 2^maxDepth / gLens[0]) per axis
 * (construct.cpp:933-941) and must lie in [0, gLens).  glens NULL = unit cube. */
int dgx_points_to_octree_glens(const double *xyz, size_t n, const double *glens, unsigned max_depth,
                               unsigned max_num_pts, int curve, int on_device, dgx_octree *out);
/* ot::balanceOctree on one rank (octUtils.h:692-694; balance.cpp:31-64):
 * 2:1 balance across faces, edges and corners.  inc_corner is accepted for
 * signature parity; in 3-D the reference always balances across corners
 * (TreeNode.cpp:1075 "dim == 3 || incCorner"). */
int dgx_balance_octree(dgx_octree in, int inc_corner, dgx_octree *out);
/* par::sampleSort for TreeNodes on one rank (parUtils.h:424-425 ->
 * omp_par::merge_sort): import n TreeNodes and sort them on the curve. */
int dgx_octree_import(const uint32_t *treenodes, size_t n, unsigned max_depth, int curve, int on_device, int sort,
                      dgx_octree *out);
int dgx_octree_size(dgx_octree o, size_t *n);
int dgx_octree_max_depth(dgx_octree o, unsigned *max_depth);
/* Copy the octants out as TreeNodes (n*7 uint32, weight 1, dim 3). */
int dgx_octree_export(dgx_octree o, uint32_t *treenodes, int on_device);
/* Counters of the last build: textbook leaves, saturated-cell triggers,
 * replays, replayed leaves (points2Octree) or levels, seeds (balance). */
int dgx_octree_stats(dgx_octree o, uint64_t out[5]);
int dgx_octree_free(dgx_octree o);

/* ---- DA ----------------------------------------------------------------- */
typedef struct {
    uint64_t local_buffer_size;  /* getLocalBufferSize(): elements + pseudo-octants */
    uint64_t element_size;       /* getElementSize() */
    uint64_t node_size;          /* m_uiNodeSize */
    uint64_t boundary_node_size; /* m_uiBoundaryNodeSize */
    uint64_t n_nodes;            /* getNodeSize() = node_size + boundary_node_size */
    uint32_t max_depth;          /* getMaxDepth(): octree depth + 1 */
    uint32_t ghost_scans;        /* node-list entries resolved by the forward ghost scan */
    /* matvec block view (no reference counterpart): element blocks, their
     * (block,node) pairs, nodes touched by more than one block, largest block */
    uint64_t matvec_blocks;
    uint64_t block_node_pairs;
    uint64_t shared_nodes;
    uint64_t shared_pairs;
    uint64_t max_block_nodes;
} dgx_da_info;

/* ot::DA::DA(in, MPI_COMM_SELF, MPI_COMM_SELF, tol, compressLut=false)
 * (include/oda/oda.h:622-624; oda.cpp:250-297).  The octree is not consumed. */
int dgx_da_create(dgx_octree octree, dgx_da *out);
int dgx_da_get_info(dgx_da da, dgx_da_info *info);
/* Host copies of the DA maps; any pointer may be NULL.
 *   nlist: 8*element_size local-buffer indices (m_uiNlist, DA::getNodeIndices)
 *   masks: element_size hanging masks (DA::getHangingNodeIndex)
 *   oct_levels: local_buffer_size flags (m_ucpOctLevels: level|BOUNDARY|NODE)
 *   local_to_node: local index -> compact node id, 0xffffffff if not a node
 *     (the vecGetBuffer compaction, oda.cpp:863-882) */
int dgx_da_export(dgx_da da, uint32_t *nlist, uint8_t *masks, uint8_t *oct_levels, uint32_t *local_to_node);
/* Compressed element-to-node maps in the reference's form (ot::DA with
 * compressLut = true; odaBuildNlist.cpp:3120-3262), encoded on the device,
 * host copies (any array pointer may be NULL):
 *   rem: 8*element_size remainders (m_ucpLutRemainders), offset % 256 of the
 *     element's sorted node indices (first offset: element index - smallest);
 *   qmask: element_size quotient masks (m_ucpLutMasks[2i]);
 *   order: element_size sort orders (m_ucpSortOrders: XYZ..ZYX, or the child
 *     number of a hanging element);
 *   quot: the nonzero quotients (unsigned short)(offset >> 8) in element order
 *     (m_uspLutQuotients); *n_quot receives their count (call with quot NULL to
 *     size the buffer; quot_cap below the count is an error). */
int dgx_da_compress_maps(dgx_da da, uint8_t *rem, uint8_t *qmask, uint8_t *order, uint16_t *quot, size_t quot_cap,
                         size_t *n_quot);
/* DA::getNodeIndices on compressed maps (include/oda/oda.tcc:1304-3950) for
 * every element, on the device: nlist (host, 8*element_size) receives the
 * decoded local-buffer indices.  With the Morton curve they equal
 * dgx_da_export's nlist.  With the Hilbert curve the reference's format does
 * not round-trip (its fixed unsort patterns assume Morton order, and the
 * 16-bit quotient of a negative first offset loses bits); this decode then
 * returns what the reference's getNodeIndices returns. */
int dgx_da_decode_maps(dgx_da da, const uint8_t *rem, const uint8_t *qmask, const uint8_t *order,
                       const uint16_t *quot, size_t n_quot, uint32_t *nlist);
/* Anchors of the local-buffer octants (3 x uint32 per octant, interleaved)
 * in the embedded 2^(maxDepth+1) domain; with oct_levels they give the
 * octant of DA::getChildNumber / getCurrentOffset (oda.tcc:1273-1300). */
int dgx_da_export_anchors(dgx_da da, uint32_t *xyz);
/* Dirichlet flags per local-buffer octant, as assignBoundaryFlags (odaUtils.cpp:
 * 2464-2812) computes them and the reference's Dirichlet drivers copy them into
 * a DAMG level's suppressedDOF (omgJac.C:553-574): for an octant that carries a
 * node, the number of axes on which the node lies on the boundary of the
 * domain (TreeNode::FACE_BDY 1, EDGE_BDY 2, CORNER_BDY 3; 0 inside); 0 for an
 * octant that carries no node.  flags holds local_buffer_size bytes (host). */
int dgx_da_boundary_flags(dgx_da da, uint8_t *flags);
/* y = (lap_fac * K + mass_fac * M) x on compact node vectors (device pointers
 * of length n_nodes).  Jacobian1MatMult (examples/src/backend/odaJac.C:
 * 297-366) with type-2 hanging stencils; lapFac/massFac as read from the
 * options -lapMatPropFac / -massMatPropFac (odaJac.C:313-317). */
int dgx_da_matvec(dgx_da da, const double *x, double *y, double lap_fac, double mass_fac);
int dgx_da_matvec_f32(dgx_da da, const float *x, float *y, float lap_fac, float mass_fac);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dgx_da_matvec with the shared nodes' partial sums in a
 * static device array of 64 MiB (8,388,608 doubles, one per GPU, held until
 * the process ends) instead of the DA's own buffer; it calls no allocation
 * routine.  Fails with an error when the DA has more partial slots than the
 * array holds (dgx_da_info.shared_pairs > 8,388,608).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int dgx_da_matvec_noalloc(dgx_da da, const double *x, double *y, double lap_fac, double mass_fac);
/* Jacobian1MatGetDiagonal (odaJac.C:160-227) on a compact node vector. */
int dgx_da_diagonal(dgx_da da, double *diag, double lap_fac, double mass_fac);
int dgx_da_free(dgx_da da);

/* Homogeneous-Dirichlet operator: DirichletJacobianMatMult (examples/src/
 * backend/omgJac.C:958-1057): rows of nodes on the boundary of the unit cube
 * are the identity, boundary columns are dropped. */
int dgx_da_matvec_dirichlet(dgx_da da, const double *x, double *y, double lap_fac, double mass_fac);

/* ---- multigrid ---------------------------------------------------------- */
typedef struct dgx_mg_s *dgx_mg;

/* DAMGCreateAndSetDA + DAMGSetKSP (include/omg/omg.h:470-528; src/omg/
 * blockDiag.cpp:780-1047, 1312-2190) on one rank: nlevels levels built by
 * coarsenOctree + balanceOctree from the finest octree (stopping early below
 * 500 octants, blockDiag.cpp:1417-1420; only a coarsest level of more than
 * 40,000 nodes, too costly for the direct solver, makes it coarsen further),
 * a DA per level, interpolation P and restriction R = P^T between levels,
 * Jacobi diagonals and the coarsest-level sparse FP64 direct solver (nested
 * dissection + multifrontal factorisation on the device, in place of the
 * reference's PETSc LU, bin/optionsFBM:57).  The octree is not consumed. */
int dgx_mg_create(dgx_octree finest, int nlevels, double lap_fac, double mass_fac, dgx_mg *out);
/* The operator factors of DAMGSetKSP's Jacobian (-lapFac lap + massFac I,
 * omgJac.C:958-1057): recomputes the level diagonals and the coarse solver. */
int dgx_mg_set_operator(dgx_mg mg, double lap_fac, double mass_fac);
/* Number of levels actually built and, per level (finest first), the number of
 * octants and of nodes (arrays of at least 32 entries, may be NULL). */
int dgx_mg_info(dgx_mg mg, int *nlevels, uint64_t *octants, uint64_t *nodes);
/* The DA of level `level` (0 = finest; owned by mg, do not free). */
int dgx_mg_level_da(dgx_mg mg, int level, dgx_da *da);
/* The finest level's DA (owned by mg; do not free). */
int dgx_mg_finest_da(dgx_mg mg, dgx_da *da);
/* coarse = R fine: restrictMatVecType1 (src/omg/restrictMatVec.cpp:486-619)
 * from level `level` to level + 1 (0 = finest), with the Dirichlet boundary
 * nodes of both levels suppressed as DAMGCreateSuppressedDOFs marks them
 * (ITLB_SET_VALUE_SUPPRESSED_DOFS, :71-82).  R holds, for every fine node i
 * and coarse node j, the value at i of the coarse basis function of j (the
 * trilinear interpolation with the coarse element's hanging-node constraints,
 * which is what RmatType1/RmatType2Stencils.inp tabulate), each pair once.
 * Device pointers to compact node vectors of the two levels. */
int dgx_mg_restrict(dgx_mg mg, int level, const double *fine, double *coarse);
/* fine = P coarse = R^T coarse: prolongMatVecType1 (src/omg/prolongMatVec.cpp:
 * 154-292) from level + 1 to level. */
int dgx_mg_prolong(dgx_mg mg, int level, const double *coarse, double *fine);
/* Load vector b = M f of a nodal f with boundary entries set to zero. */
int dgx_mg_rhs(dgx_mg mg, const double *f, double *b);

typedef struct {
    int fp32_vcycle; /* 1: V-cycle in FP32 inside FP64 CG; 0: all FP64 */
    int pre, post;   /* Jacobi sweeps before / after the coarse correction */
    double omega;    /* Jacobi damping */
    double rtol;     /* stop at ||b - A u|| <= rtol ||b|| */
    int maxit;
    int use_graph;   /* 1: replay captured HIP graphs of the CG step and V-cycle */
} dgx_solve_opts;
typedef struct {
    int iterations;
    double rel_residual; /* recomputed in FP64 after the last iteration */
    double seconds;
} dgx_solve_stats;
/* DAMGSolve (blockDiag.cpp:1122-1170): CG preconditioned by one V-cycle per
 * iteration on device vectors b, u of n_nodes entries (finest level). */
int dgx_mg_solve(dgx_mg mg, const double *b, double *u, const dgx_solve_opts *opts, dgx_solve_stats *stats);
int dgx_mg_free(dgx_mg mg);

/* Coarsest-level direct solver of a hierarchy (the reference's coarse PCLU,
 * bin/optionsFBM:57): out[0] interior nodes, [1] fronts, [2] levels of the
 * elimination tree, [3] largest front, [4] largest separator, [5] doubles in
 * the frontal matrices. */
int dgx_mg_coarse_info(dgx_mg mg, uint64_t out[6]);

/* Symbolic analysis of the coarsest-level solver on host arrays (no device):
 * nested dissection of the interior-node graph of n nodes and n_elem elements
 * (conn: 8 x n_elem node ids, conn[j*n_elem + e]; meta: mask | child << 8 |
 * level << 11 per element; bdy: 1 for boundary nodes; xyz: 3 coordinates per
 * node), frontal index sets and the places of the matrix entries in the fronts.
 * Parts of at most `leaf` interior nodes are not split (0: 128); at most
 * `dense` interior nodes make one front.  For inspection and tests. */
typedef struct dgx_direct_plan_s *dgx_direct_plan;
typedef struct {
    uint64_t interior, fronts, links, levels, entries, idx_len, inv_len, fmap_len, a_nnz, a_src_len, upd_len;
    uint32_t max_m, max_s;
} dgx_direct_sizes;
int dgx_direct_plan_host(uint32_t n, uint64_t n_elem, const uint32_t *conn, const uint16_t *meta, const uint8_t *bdy,
                         const uint32_t *xyz, uint32_t leaf, uint32_t dense, dgx_direct_plan *out,
                         dgx_direct_sizes *sz);
/* Copies the plan out (any pointer may be NULL):
 *   front9: per front (in elimination order) offset of its dense matrix, size m,
 *     pivots s, offset of its index set in idx, of its update vector, of its
 *     forward values, first link, number of children, height in the tree;
 *   link3: per link child front, offset in inv (m of the parent entries: the
 *     child's update index of each parent position or -1), offset in fmap
 *     (parent position of each child update index);
 *   idx: node ids of the fronts' index sets (pivots first);
 *   a_off/a_src/a_dst: per matrix entry of the interior rows (columns sorted),
 *     its element-matrix contributions e*64 + k*8 + j (row j, column k of the
 *     element's child-frame matrix) and its offset in the fronts;
 *   inode: interior index -> node id. */
int dgx_direct_plan_export(dgx_direct_plan h, uint64_t *front9, uint64_t *link3, int32_t *inv, uint32_t *fmap,
                           uint32_t *idx, uint32_t *a_off, uint32_t *a_src, uint64_t *a_dst, uint32_t *inode);
int dgx_direct_plan_free(dgx_direct_plan h);

/* ---- request server ----------------------------------------------------- */
/* No counterpart in the reference (its drivers solve one problem per run):
 * a server for streams of Poisson problems on point clouds.  Every request is
 * points2Octree + balanceOctree + DAMGCreateAndSetDA + DAMGSolve of
 * -lap u + u = f (homogeneous Dirichlet, f nodal N(0,1) from rhs_seed).
 * `slots` requests run at once, each on its own stream, and a request is
 * admitted when a slot is free and its memory estimate fits hbm_budget next to
 * the running requests and the resident problems; finished problems stay
 * resident (LRU, keyed by a hash of the points and parameters) when
 * keep_resident is set, so a repeated request only re-solves. */
typedef struct dgx_server_s *dgx_server;
typedef struct {
    unsigned slots;        /* 1..16 */
    uint64_t hbm_budget;   /* bytes */
    int nlevels;           /* multigrid levels requested */
    int fp32_vcycle;       /* 1: FP32 V-cycle inside FP64 CG */
    double rtol;
    int maxit;
    int keep_resident;
    int return_solution;   /* 1: keep a host copy of each solution for dgx_server_wait */
} dgx_server_opts;
typedef struct {
    uint64_t id;
    int status;            /* 0 ok */
    int iterations;
    double rel_residual;
    uint64_t octants, nodes;
    int levels;
    int reused;            /* 1: the hierarchy was resident */
    double t_submit, t_start, t_done; /* seconds on the server clock */
    uint64_t peak_bytes;     /* peak device memory the request held */
    uint64_t resident_bytes; /* device memory of the problem kept resident (0 if not kept) */
    double s_octree, s_hierarchy, s_solve; /* seconds in the octree build, the MG set-up and the solve */
} dgx_served;
int dgx_server_create(const dgx_server_opts *opts, dgx_server *out);
/* Copies the n points; returns at once with the request id. */
int dgx_server_submit(dgx_server sv, const double *xyz, size_t n, unsigned max_depth, unsigned max_num_pts,
                      uint64_t rhs_seed, uint64_t *id);
/* Blocks until request id is done; x (host, x_len >= nodes) receives the
 * solution when return_solution is set, or may be NULL. */
int dgx_server_wait(dgx_server sv, uint64_t id, dgx_served *res, double *x, size_t x_len);
/* Blocks until request id is done and reports it without consuming it. */
int dgx_server_peek(dgx_server sv, uint64_t id, dgx_served *res);
double dgx_server_clock(dgx_server sv);
/* A result can be waited for once: a second dgx_server_wait or dgx_server_peek
 * on a consumed id returns an error instead of blocking. */
/* Admission state (no reference counterpart): the effective budget, the bytes
 * reserved by running requests' estimates plus resident problems, the
 * resident problems and their measured bytes, queued and running requests,
 * and the number of evictions so far. */
typedef struct {
    uint64_t budget, reserved, resident, resident_bytes, queued, running, evictions;
} dgx_server_stats;
int dgx_server_get_stats(dgx_server sv, dgx_server_stats *st);
/* Fault injection for tests of the error paths: DGX_SERVER_FAULT_RESOLVE makes
 * the next re-solve of a resident problem fail; the problem is then dropped
 * from the cache and its bytes are returned to the budget. */
enum { DGX_SERVER_FAULT_RESOLVE = 1 };
int dgx_server_inject_fault(dgx_server sv, int kind);
/* Finishes the queued requests, then frees everything. */
int dgx_server_free(dgx_server sv);

/* ---- on-disk formats (host only; need no dgx_init) ---------------------- */
/* Readers use two calls: with a NULL buffer (or cap below the count) they only
 * set *n; then call again with a buffer of cap >= *n entries.  Octants are
 * ot::TreeNode rows of 7 x u32 (x, y, z, level|flags, weight, dim, maxDepth). */
/* ot::readPtsFromFile (src/octFunctions.cpp:319-348): u32 N, then 3N doubles.
 * An empty file holds no points (what writePtsToFile writes for none). */
int dgx_read_pts(const char *path, double *pts, size_t cap, size_t *n);
/* ot::writePtsToFile (src/octFunctions.cpp:380-395): nothing when n == 0. */
int dgx_write_pts(const char *path, const double *pts, size_t n);
/* ot::readDataPtsFromFile (src/octFunctions.cpp:352-378): u32 N, 3N doubles of
 * points, N doubles of values. */
int dgx_read_data_pts(const char *path, double *pts, double *vals, size_t cap, size_t *n);
/* ot::writeDataPtsToFile (src/octFunctions.cpp:398-412): writes N and the
 * points; like the reference it leaves the values out (vals is ignored). */
int dgx_write_data_pts(const char *path, const double *pts, const double *vals, size_t n);
/* ot::readNodesFromFile (src/octFunctions.cpp:433-453): text "dim maxDepth N"
 * then "x y z level" per octant; weight 1, no flags (TreeNode.cpp:650-662). */
int dgx_read_nodes(const char *path, uint32_t *tn, size_t cap, size_t *n);
/* ot::writeNodesToFile (src/octFunctions.cpp:469-486); all octants must share
 * dim and maxDepth; the level is written without flags. */
int dgx_write_nodes(const char *path, const uint32_t *tn, size_t n);
/* ot::readNodesFromFile_binary (src/octFunctions.cpp:415-431): 28-byte
 * records.  Returns the N records of the file; the reference's feof loop
 * appends a copy of the last one. */
int dgx_read_nodes_binary(const char *path, uint32_t *tn, size_t cap, size_t *n);
/* ot::writeNodesToFile_binary (src/octFunctions.cpp:455-467). */
int dgx_write_nodes_binary(const char *path, const uint32_t *tn, size_t n);
/* treeNodesTovtk (src/treenode2vtk.cpp:4-134): sorts tn in place along the
 * curve (as the reference's std::sort does), then writes
 * <name>_H_<rank>.vtk (Hilbert) or <name>_M_<rank>.vtk (Morton). */
int dgx_write_vtk(uint32_t *tn, size_t n, int rank, const char *name, int hsorted, int curve);

/* ---- inputs ------------------------------------------------------------- */
/* Deterministic synthetic point clouds (host memory, n*3 doubles):
 *   dist 0: i.i.d. uniform [0,1)^3
 *   dist 1: each coordinate exp(N(-1, 0.5)) clipped to [0,1)
 *   dist 2: mixture of n_gauss isotropic Gaussians, centres U[lo,hi]^3,
 *           sigma sig_lo (or log-uniform in [sig_lo, sig_hi] when sig_hi > 0),
 *           equal weights (or Dirichlet(1) weights when dirichlet != 0), clipped.
 * SplitMix64 keyed by (seed, stream, counter); Box-Muller normals. */
/* The SFC tables as used by the library (24 x 8 each): rank of Morton child
 * m under rotation r, Morton child of rank k, rotation entered in Morton
 * child m (HILBERT_TABLE / rotations, src/hcurvedata.cpp:329-546; identity
 * for Morton).  Any pointer may be NULL. */
int dgx_sfc_tables(int curve, uint8_t *rank192, uint8_t *morton192, uint8_t *crot192);
int dgx_gen_points(int dist, uint64_t seed, size_t n, int n_gauss, double lo, double hi, double sig_lo, double sig_hi,
                   int dirichlet, double *out);

#ifdef __cplusplus
}
#endif
#endif
