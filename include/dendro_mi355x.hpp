// dendro_mi355x.hpp -- C++ facade with the reference's one-rank par / ot /
// DAMG entry points, implemented on the C ABI of dendro_mi355x.h.
//
// The names, argument lists and argument meanings follow the reference:
//   ot::TreeNode                      include/TreeNode.h:39-472
//   ot::points2Octree                 include/octUtils.h:633-634
//   ot::balanceOctree                 include/octUtils.h:692-694
//   par::sampleSort                   include/parUtils.h:424-425
//   ot::DA                            include/oda/oda.h:622-1460
//   DAMG_Initialize .. DAMGDestroy    include/omg/omg.h:419-608
// Differences, all forced by running on one GPU without MPI or PETSc:
//   * only one rank: a communicator with more than one rank is rejected
//     (when mpi.h is included before this header; otherwise MPI_Comm is a
//     placeholder type and MPI_COMM_SELF / MPI_COMM_WORLD are 0);
//   * 3-D only (dim must be 3), maxDepth <= 30;
//   * the DAMG operator is the homogeneous-Dirichlet -lapFac lap + massFac I
//     of the reference drivers (omgJac.C:958-1057); DAMGSetKSP takes the two
//     factors and a right-hand-side callback that fills a host vector of
//     getNodeSize() entries, and x / b are host vectors (no PETSc Vec / Mat);
//   * errors throw dgx::Error (the reference asserts); success returns 1 like
//     the reference's non-PETSc build (PROF_*_END, octUtils.h:408-445).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "dendro_mi355x.h"

#ifdef MPI_VERSION
#define DGX_FACADE_HAVE_MPI 1
#else
#define DGX_FACADE_HAVE_MPI 0
typedef int MPI_Comm;
#ifndef MPI_COMM_SELF
#define MPI_COMM_SELF 0
#endif
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD 0
#endif
#endif

namespace ot {
struct TreeNode;
}

namespace dgx {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

inline void check(int rc, const char *what) {
    if (rc != 0) throw Error(std::string(what) + ": " + dgx_last_error());
}

inline void require_one_rank(MPI_Comm comm) {
#if DGX_FACADE_HAVE_MPI
    int np = 1;
    MPI_Comm_size(comm, &np);
    if (np != 1) throw Error("this build runs on one rank per GPU (communicator has " + std::to_string(np) + ")");
#else
    (void)comm;
#endif
}

// Curve used by TreeNode ordering and every build call; Hilbert matches a
// reference build with HILBERT_ORDERING (its default).
inline int &curve() {
    static int c = DGX_CURVE_HILBERT;
    return c;
}

// Device selection: DGX_DEVICE (default 0); done once per process.
inline void ensure_init() {
    static bool done = false;
    if (done) return;
    const char *e = std::getenv("DGX_DEVICE");
    check(dgx_init(e ? std::atoi(e) : 0), "dgx_init");
    done = true;
}

namespace detail {
struct Tables {
    uint8_t rank[192], crot[192];
};
inline const Tables &tables(int c) {
    static Tables t[2];
    static bool ok[2] = {false, false};
    if (!ok[c]) {
        check(dgx_sfc_tables(c, t[c].rank, nullptr, t[c].crot), "dgx_sfc_tables");
        ok[c] = true;
    }
    return t[c];
}
// 128-bit SFC key: the curve rank of every level's child digit, most
// significant first, then the level; unsigned order == TreeNode::operator<.
inline std::pair<uint64_t, uint64_t> key(uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    const Tables &T = tables(curve());
    uint64_t hi = 0, lo = 0;
    uint32_t rot = 0;
    for (uint32_t t = 1; t <= lev; ++t) {
        const uint32_t b = md - t;
        const uint32_t m = (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
        const uint64_t r = T.rank[rot * 8 + m];
        rot = T.crot[rot * 8 + m];
        if (t <= 21) hi |= r << (63 - 3 * t);
        else lo |= r << (63 - 3 * (t - 21));
    }
    return {hi, lo | lev};
}

class Octree {
  public:
    Octree() = default;
    explicit Octree(dgx_octree h) : h_(h) {}
    Octree(const Octree &) = delete;
    Octree &operator=(const Octree &) = delete;
    Octree(Octree &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    ~Octree() {
        if (h_) dgx_octree_free(h_);
    }
    dgx_octree get() const { return h_; }
    dgx_octree *out() { return &h_; }

  private:
    dgx_octree h_ = nullptr;
};

// Host TreeNode vector <-> device octree.
Octree upload(const std::vector<ot::TreeNode> &v, unsigned maxDepth, bool sort);
void download(const Octree &o, std::vector<ot::TreeNode> &v);
}  // namespace detail
}  // namespace dgx

namespace ot {

// 7 x uint32, the reference layout (TreeNode.h:42).
struct TreeNode {
    enum { MAX_LEVEL = 31, NODE = 128, BOUNDARY = 64, DEP_ELEM = 32 };  // TreeNode.h:79-81
    uint32_t m_uiX = 0, m_uiY = 0, m_uiZ = 0, m_uiLevel = 0, m_uiWeight = 1, m_uiDim = 3, m_uiMaxDepth = 0;

    TreeNode() = default;
    TreeNode(unsigned dim, unsigned maxDepth) : m_uiDim(dim), m_uiMaxDepth(maxDepth) {}  // the root
    TreeNode(unsigned x, unsigned y, unsigned z, unsigned level, unsigned dim, unsigned maxDepth)
        : m_uiX(x), m_uiY(y), m_uiZ(z), m_uiLevel(level), m_uiDim(dim), m_uiMaxDepth(maxDepth) {}

    unsigned getX() const { return m_uiX; }
    unsigned getY() const { return m_uiY; }
    unsigned getZ() const { return m_uiZ; }
    unsigned getLevel() const { return m_uiLevel & MAX_LEVEL; }
    unsigned getFlag() const { return m_uiLevel; }
    unsigned getWeight() const { return m_uiWeight; }
    unsigned getDim() const { return m_uiDim; }
    unsigned getMaxDepth() const { return m_uiMaxDepth; }
    void setWeight(unsigned w) { m_uiWeight = w; }
    bool operator==(const TreeNode &o) const {  // TreeNode.tcc:93-107
        return m_uiX == o.m_uiX && m_uiY == o.m_uiY && m_uiZ == o.m_uiZ && getLevel() == o.getLevel();
    }
    bool operator!=(const TreeNode &o) const { return !(*this == o); }
    // SFC order with ancestors first (TreeNode.tcc:122-238)
    bool operator<(const TreeNode &o) const {
        return dgx::detail::key(m_uiX, m_uiY, m_uiZ, getLevel(), m_uiMaxDepth) <
               dgx::detail::key(o.m_uiX, o.m_uiY, o.m_uiZ, o.getLevel(), o.m_uiMaxDepth);
    }
    bool operator>(const TreeNode &o) const { return o < *this; }
    bool operator<=(const TreeNode &o) const { return !(o < *this); }
    bool operator>=(const TreeNode &o) const { return !(*this < o); }
};
static_assert(sizeof(TreeNode) == 28, "TreeNode must keep the reference's 28-byte layout");

// _InitializeHcurve (hcurvedata.h:45): one-time set-up of the curve tables;
// here it selects the device and creates the library stream.
inline void _InitializeHcurve(int dim) {
    if (dim != 3) throw dgx::Error("_InitializeHcurve: only dim == 3 is supported");
    dgx::ensure_init();
}

// ot::points2Octree (octUtils.h:633-634; construct.cpp:759-974): pts holds
// 3 coordinates per point in [0, gLens); pts is cleared (construct.cpp:944).
inline int points2Octree(std::vector<double> &pts, double *gLens, std::vector<TreeNode> &nodes, unsigned dim,
                         unsigned maxDepth, unsigned maxNumPts, MPI_Comm comm, double tol = 0.1) {
    (void)tol;  // partition tolerance: one rank
    dgx::require_one_rank(comm);
    if (dim != 3) throw dgx::Error("points2Octree: only dim == 3 is supported");
    dgx::ensure_init();
    dgx::detail::Octree o;
    dgx::check(dgx_points_to_octree_glens(pts.data(), pts.size() / 3, gLens, maxDepth, maxNumPts, dgx::curve(), 0,
                                          o.out()),
               "points2Octree");
    dgx::detail::download(o, nodes);
    pts.clear();
    return 1;
}

// ot::balanceOctree (octUtils.h:692-694; balance.cpp:31-64): 2:1 balance of a
// sorted, complete, linear octree; in is cleared (balance.cpp:61).
inline int balanceOctree(std::vector<TreeNode> &in, std::vector<TreeNode> &out, unsigned dim, unsigned maxDepth,
                         bool incCorner, MPI_Comm comm, MPI_Comm *newCommPtr = nullptr, bool *iAmActive = nullptr) {
    dgx::require_one_rank(comm);
    if (dim != 3) throw dgx::Error("balanceOctree: only dim == 3 is supported");
    dgx::ensure_init();
    if (newCommPtr) *newCommPtr = comm;
    if (iAmActive) *iAmActive = !in.empty();
    dgx::detail::Octree src = dgx::detail::upload(in, maxDepth, false);
    dgx::detail::Octree dst;
    dgx::check(dgx_balance_octree(src.get(), incCorner ? 1 : 0, dst.out()), "balanceOctree");
    dgx::detail::download(dst, out);
    in.clear();
    return 1;
}

// ---- on-disk formats (octFunctions.cpp:319-488); host only ----------------
namespace detail_io {
template <typename T, typename Read>
inline void read_all(const char *what, std::vector<T> &v, size_t width, Read rd) {
    size_t n = 0;
    dgx::check(rd(nullptr, 0, &n), what);
    v.resize(width * n);
    if (n) dgx::check(rd(v.data(), n, &n), what);
}
}  // namespace detail_io

// ot::readPtsFromFile (octFunctions.cpp:319-348): pts receives 3N coordinates.
inline int readPtsFromFile(const char *filename, std::vector<double> &pts) {
    detail_io::read_all("readPtsFromFile", pts, 3, [&](double *p, size_t cap, size_t *n) {
        return dgx_read_pts(filename, p, cap, n);
    });
    return 1;
}
// ot::readDataPtsFromFile (octFunctions.cpp:352-378).
inline int readDataPtsFromFile(const char *filename, std::vector<double> &pts, std::vector<double> &ptVals) {
    size_t n = 0;
    dgx::check(dgx_read_data_pts(filename, nullptr, nullptr, 0, &n), "readDataPtsFromFile");
    pts.resize(3 * n);
    ptVals.resize(n);
    if (n) dgx::check(dgx_read_data_pts(filename, pts.data(), ptVals.data(), n, &n), "readDataPtsFromFile");
    return 1;
}
// ot::writePtsToFile (octFunctions.cpp:380-395).
inline int writePtsToFile(const char *filename, std::vector<double> &pts) {
    dgx::check(dgx_write_pts(filename, pts.data(), pts.size() / 3), "writePtsToFile");
    return 1;
}
// ot::writeDataPtsToFile (octFunctions.cpp:398-412): N and the points; the
// values are left out, as in the reference.
inline int writeDataPtsToFile(const char *filename, std::vector<double> &pts, std::vector<double> &data) {
    dgx::check(dgx_write_data_pts(filename, pts.data(), data.data(), pts.size() / 3), "writeDataPtsToFile");
    return 1;
}
// ot::readNodesFromFile (octFunctions.cpp:433-453): the ".ot" text format.
inline int readNodesFromFile(const char *filename, std::vector<TreeNode> &nodes) {
    size_t n = 0;
    dgx::check(dgx_read_nodes(filename, nullptr, 0, &n), "readNodesFromFile");
    nodes.resize(n);
    if (n) dgx::check(dgx_read_nodes(filename, reinterpret_cast<uint32_t *>(nodes.data()), n, &n), "readNodesFromFile");
    return 1;
}
// ot::readNodesFromFile_binary (octFunctions.cpp:415-431): appends the N
// records of the file (the reference's feof loop appends the last one twice)
// and returns the new size.
inline int readNodesFromFile_binary(const char *filename, std::vector<TreeNode> &nodes) {
    size_t n = 0;
    dgx::check(dgx_read_nodes_binary(filename, nullptr, 0, &n), "readNodesFromFile_binary");
    const size_t old = nodes.size();
    nodes.resize(old + n);
    if (n)
        dgx::check(dgx_read_nodes_binary(filename, reinterpret_cast<uint32_t *>(nodes.data() + old), n, &n),
                   "readNodesFromFile_binary");
    return (int)nodes.size();
}
// ot::writeNodesToFile_binary (octFunctions.cpp:455-467).
inline int writeNodesToFile_binary(const char *filename, const std::vector<TreeNode> &nodes) {
    dgx::check(dgx_write_nodes_binary(filename, reinterpret_cast<const uint32_t *>(nodes.data()), nodes.size()),
               "writeNodesToFile_binary");
    return 0;
}
// ot::writeNodesToFile (octFunctions.cpp:469-486).
inline int writeNodesToFile(const char *filename, const std::vector<TreeNode> &nodes) {
    dgx::check(dgx_write_nodes(filename, reinterpret_cast<const uint32_t *>(nodes.data()), nodes.size()),
               "writeNodesToFile");
    return 1;
}

namespace DA_FLAGS {
// Element loops (oda.h); on one rank every element is writable and
// independent, DEPENDENT loops are empty.
enum loopType { ALL, WRITABLE, INDEPENDENT, DEPENDENT, W_DEPENDENT };
}  // namespace DA_FLAGS

// ot::DA on one rank (oda.h:622-1460).  The element-to-node maps and sizes are
// those of the reference with compressLut == false; the octants stay on the
// device, host copies of the maps are made on first use.
class DA {
  public:
    DA(std::vector<TreeNode> &in, MPI_Comm comm, MPI_Comm inputActiveComm, double tol, bool compressLut = false,
       const std::vector<TreeNode> *blocksPtr = nullptr, bool *iAmActive = nullptr) {
        (void)inputActiveComm;
        (void)tol;
        (void)blocksPtr;
        // compressLut: the element-to-node maps are held on the host in the
        // reference's compressed form and decoded per element by
        // getNodeIndices.  Only with Morton order: with Hilbert order the
        // reference's format does not round-trip (dgx_da_decode_maps), so the
        // plain maps are kept.
        compress_ = compressLut && dgx::curve() == DGX_CURVE_MORTON;
        dgx::require_one_rank(comm);
        dgx::ensure_init();
        if (iAmActive) *iAmActive = !in.empty();
        unsigned md = in.empty() ? 0 : in[0].getMaxDepth();
        dgx::detail::Octree o = dgx::detail::upload(in, md, false);
        dgx::check(dgx_da_create(o.get(), &h_), "DA");
        owned_ = true;
        info();
        in.clear();  // the DA consumes its input (oda.cpp:250-297)
    }
    // Wraps a DA owned by someone else (the finest level of a DAMG).
    explicit DA(dgx_da borrowed) : h_(borrowed) { info(); }
    DA(const DA &) = delete;
    DA &operator=(const DA &) = delete;
    ~DA() {
        if (owned_ && h_) dgx_da_free(h_);
    }

    unsigned getNodeSize() const { return (unsigned)inf_.n_nodes; }  // oda.tcc:156-158
    unsigned getBoundaryNodeSize() const { return (unsigned)inf_.boundary_node_size; }
    unsigned getElementSize() const { return (unsigned)inf_.element_size; }
    unsigned getLocalBufferSize() const { return (unsigned)inf_.local_buffer_size; }
    unsigned getMaxDepth() const { return inf_.max_depth; }
    unsigned getElementBegin() const { return 0; }
    unsigned getElementEnd() const { return (unsigned)inf_.element_size; }
    unsigned getIndependentSize() const { return (unsigned)inf_.element_size; }

    // Element iteration (oda.tcc:263-660) in local-buffer order.
    template <DA_FLAGS::loopType type>
    unsigned init() {
        maps();
        cur_ = (type == DA_FLAGS::DEPENDENT || type == DA_FLAGS::W_DEPENDENT) ? getElementEnd() : 0;
        return cur_;
    }
    template <DA_FLAGS::loopType type>
    unsigned end() const {
        return getElementEnd();
    }
    template <DA_FLAGS::loopType type>
    unsigned next() {
        return ++cur_;
    }
    unsigned curr() const { return cur_; }
    // Local-buffer indices of the current element's 8 vertices (m_uiNlist),
    // decoded from the compressed maps when the DA keeps them (oda.tcc:1304-3950).
    int getNodeIndices(unsigned *nodes) {
        maps();
        if (compress_) {
            decode(cur_, nodes);
            return 1;
        }
        std::memcpy(nodes, &nlist_[8 * (size_t)cur_], 8 * sizeof(unsigned));
        return 1;
    }
    bool isLutCompressed() const { return compress_; }
    // Hanging mask of element i (m_ucpLutMasks[2i+1]).
    unsigned char getHangingNodeIndex(unsigned i) {
        maps();
        return masks_[i];
    }
    // Whether vertex i of the current element is hanging.
    bool isHanging(unsigned i) {
        maps();
        return (masks_[cur_] >> i) & 1u;
    }
    // Level (embedded, i.e. octree level + 1) and flags of local octant i.
    unsigned char getLevel(unsigned i) {
        maps();
        return (unsigned char)(levels_[i] & TreeNode::MAX_LEVEL);
    }
    unsigned char getFlag(unsigned i) {
        maps();
        return levels_[i];
    }
    // Morton child number of the current element (oda.tcc:1273-1300).
    unsigned char getChildNumber() {
        anchors();
        const unsigned md = getMaxDepth(), lev = getLevel(cur_);
        const unsigned b = md - lev;
        const uint32_t *a = &anchors_[3 * (size_t)cur_];
        return (unsigned char)(((a[0] >> b) & 1u) | (((a[1] >> b) & 1u) << 1) | (((a[2] >> b) & 1u) << 2));
    }
    // Local-buffer index -> compact node id (0xffffffff if the octant is no node).
    const std::vector<uint32_t> &localToNode() {
        maps();
        return l2n_;
    }

    // Vectors (oda.cpp:697-1020): compact nodal vectors have getNodeSize()
    // entries, ghosted ones getLocalBufferSize(); elemental ones one entry per
    // element (local buffer when ghosted).
    template <typename T>
    int createVector(std::vector<T> &v, bool isElemental, bool isGhosted, unsigned dof = 1) const {
        size_t n = isElemental ? (isGhosted ? getLocalBufferSize() : getElementSize())
                               : (isGhosted ? getLocalBufferSize() : getNodeSize());
        v.assign(n * dof, T(0));
        return 1;
    }
    // Local-buffer view of a compact nodal vector (and back).
    template <typename T>
    int vecGetBuffer(std::vector<T> &in, T *&out, bool isElemental, bool isGhosted, bool isReadOnly,
                     unsigned dof = 1) {
        (void)isReadOnly;
        maps();
        const size_t n = getLocalBufferSize();
        out = new T[n * dof]();
        if (isElemental) {
            const size_t ne = isGhosted ? n : getElementSize();
            for (size_t i = 0; i < ne * dof && i < in.size(); ++i) out[i] = in[i];
            return 1;
        }
        for (size_t i = 0; i < n; ++i)
            if (l2n_[i] != 0xffffffffu)
                for (unsigned d = 0; d < dof; ++d) out[i * dof + d] = in[(size_t)l2n_[i] * dof + d];
        return 1;
    }
    template <typename T>
    int vecRestoreBuffer(std::vector<T> &in, T *out, bool isElemental, bool isGhosted, bool isReadOnly,
                         unsigned dof = 1) {
        if (!isReadOnly) {
            if (isElemental) {
                const size_t ne = isGhosted ? getLocalBufferSize() : getElementSize();
                for (size_t i = 0; i < ne * dof && i < in.size(); ++i) in[i] = out[i];
            } else {
                for (size_t i = 0; i < l2n_.size(); ++i)
                    if (l2n_[i] != 0xffffffffu)
                        for (unsigned d = 0; d < dof; ++d) in[(size_t)l2n_[i] * dof + d] = out[i * dof + d];
            }
        }
        delete[] out;
        return 1;
    }

    // Jacobian1MatMult (odaJac.C:275-366) on the device: y = (lapFac K +
    // massFac M) x, compact nodal device vectors.
    void matvec(const double *x_dev, double *y_dev, double lapFac, double massFac) {
        dgx::check(dgx_da_matvec(h_, x_dev, y_dev, lapFac, massFac), "DA::matvec");
    }
    // Same on host vectors of getNodeSize() entries (copies in and out).
    void matvec(const std::vector<double> &x, std::vector<double> &y, double lapFac, double massFac) {
        const size_t n = getNodeSize(), bytes = n * sizeof(double);
        if (x.size() != n) throw dgx::Error("DA::matvec: x must have getNodeSize() entries");
        void *dx = nullptr, *dy = nullptr;
        dgx::check(dgx_device_malloc(&dx, bytes ? bytes : 8), "malloc");
        dgx::check(dgx_device_malloc(&dy, bytes ? bytes : 8), "malloc");
        int rc = dgx_memcpy_htod(dx, x.data(), bytes);
        if (!rc) rc = dgx_da_matvec(h_, (const double *)dx, (double *)dy, lapFac, massFac);
        y.resize(n);
        if (!rc) rc = dgx_memcpy_dtoh(y.data(), dy, bytes);
        dgx_device_free(dx);
        dgx_device_free(dy);
        dgx::check(rc, "DA::matvec");
    }
    dgx_da handle() const { return h_; }

  private:
    void info() { dgx::check(dgx_da_get_info(h_, &inf_), "DA info"); }
    void maps() {
        if (have_maps_) return;
        const size_t ne = inf_.element_size;
        if (!compress_) nlist_.resize(8 * ne);
        masks_.resize(ne);
        levels_.resize(inf_.local_buffer_size);
        l2n_.resize(inf_.local_buffer_size);
        dgx::check(dgx_da_export(h_, compress_ ? nullptr : nlist_.data(), masks_.data(), levels_.data(), l2n_.data()),
                   "DA maps");
        if (compress_) {
            size_t nq = 0;
            dgx::check(dgx_da_compress_maps(h_, nullptr, nullptr, nullptr, nullptr, 0, &nq), "DA compressed maps");
            rem_.resize(8 * ne);
            qmask_.resize(ne);
            order_.resize(ne);
            quot_.resize(nq);
            dgx::check(dgx_da_compress_maps(h_, rem_.data(), qmask_.data(), order_.data(), quot_.data(), nq, &nq),
                       "DA compressed maps");
            // each element's first quotient (the reference keeps a running counter)
            qoff_.resize(ne);
            uint32_t q = 0;
            for (size_t i = 0; i < ne; ++i) {
                qoff_[i] = q;
                q += (uint32_t)__builtin_popcount(qmask_[i]);
            }
        }
        have_maps_ = true;
    }
    // DA::getNodeIndices with compressLut (oda.tcc:1304-3950): the sorted
    // indices from remainders and quotients, then unsorted by the fixed
    // pattern of the sort order, or for a hanging element by the curve order
    // of its vertex points (a hanging vertex j moved to the parent's corner j).
    void decode(unsigned i, unsigned *nodes) {
        static const unsigned char pat[6][6] = {{1, 2, 3, 4, 5, 6}, {2, 1, 3, 4, 6, 5}, {1, 4, 5, 2, 3, 6},
                                                {4, 1, 5, 2, 6, 3}, {2, 4, 6, 1, 3, 5}, {4, 2, 6, 1, 5, 3}};
        unsigned nn[8];
        uint32_t qc = qoff_[i];
        for (int j = 0; j < 8; ++j) {
            unsigned v = rem_[8 * (size_t)i + j];
            if ((qmask_[i] >> j) & 1u) v += (unsigned)quot_[qc++] << 8;
            nn[j] = j == 0 ? i - v : v + nn[j - 1];
        }
        const unsigned h = masks_[i], c = order_[i];
        if (!h) {
            nodes[0] = nn[0];
            for (int k = 0; k < 6; ++k) nodes[pat[c][k]] = nn[k + 1];
            nodes[7] = nn[7];
            return;
        }
        anchors();
        const unsigned md = getMaxDepth(), sz = 1u << (md - getLevel(i));
        const uint32_t *a = &anchors_[3 * (size_t)i];
        const unsigned p[3] = {a[0] - (c & 1u) * sz, a[1] - ((c >> 1) & 1u) * sz, a[2] - ((c >> 2) & 1u) * sz};
        TreeNode v[8];
        int idx[8];
        for (int j = 0; j < 8; ++j) {
            const bool hj = (h >> j) & 1u;
            const unsigned s = hj ? 2 * sz : sz;
            unsigned q[3];
            for (int t = 0; t < 3; ++t) q[t] = (hj ? p[t] : a[t]) + ((j >> t) & 1) * s;
            v[j] = TreeNode(q[0], q[1], q[2], md, 3, md);
            idx[j] = j;
        }
        std::sort(idx, idx + 8, [&](int x, int y) { return v[x] < v[y]; });
        for (int k = 0; k < 8; ++k) nodes[idx[k]] = nn[k];
    }
    void anchors() {
        maps();
        if (!anchors_.empty() || !inf_.local_buffer_size) return;
        anchors_.resize(3 * (size_t)inf_.local_buffer_size);
        dgx::check(dgx_da_export_anchors(h_, anchors_.data()), "DA anchors");
    }
    dgx_da h_ = nullptr;
    bool owned_ = false, have_maps_ = false, compress_ = false;
    dgx_da_info inf_{};
    unsigned cur_ = 0;
    std::vector<uint32_t> nlist_, l2n_, anchors_, qoff_;
    std::vector<uint8_t> masks_, levels_;
    std::vector<uint8_t> rem_, qmask_, order_;  // compressed maps (compressLut)
    std::vector<uint16_t> quot_;
};

// ---- DAMG (omg.h:419-608) ---------------------------------------------------
struct _p_DAMG {
    DA *da = nullptr;          // mesh of this level (the finest level only; coarser levels live on the device)
    unsigned dof = 1;
    int nlevels = 0;           // on damg[0]: the total number of levels
    int totalLevels = 0;
    MPI_Comm comm{};
    void *user = nullptr;
    std::vector<double> x, b;  // solution and right-hand side (finest level, getNodeSize() entries)
    // solver set-up (DAMGSetKSP)
    double lapFac = 1.0, massFac = 1.0;
    int (*rhs)(_p_DAMG *, std::vector<double> &) = nullptr;
    dgx_solve_opts opts{1, 2, 2, 0.8, 1e-8, 200, 1};
    dgx_solve_stats stats{};
    dgx_mg mg = nullptr;       // shared by all levels; owned by damg[0]
    int level = 0;             // level in the device hierarchy (0 = finest)
    // Dirichlet marks per local-buffer entry (dof * getLocalBufferSize() bytes),
    // set by DAMGCreateSuppressedDOFs (omg.h:427)
    unsigned char *suppressedDOF = nullptr;
    unsigned char *suppressedDOFaux = nullptr;  // no auxiliary DAs on one rank: stays NULL
};
typedef _p_DAMG *DAMG;

inline int DAMG_Initialize(MPI_Comm comm) {
    dgx::require_one_rank(comm);
    dgx::ensure_init();
    return 0;
}
inline int DAMG_Finalize() { return dgx_synchronize(); }

// DAMGCreateAndSetDA (omg.h:470; blockDiag.cpp:1312-2190): builds the
// hierarchy by coarsening and balancing the finest octree.  nlevels is
// reduced when the coarse octree has fewer than 500 octants and may grow so
// that the coarsest level fits the dense FP64 coarse solve; on return it holds
// the number of levels built.  damg[0] is the coarsest level, damg[nlevels-1]
// the finest.  finestOctree is cleared.
inline int DAMGCreateAndSetDA(MPI_Comm comm, int &nlevels, void *user, DAMG **damg,
                              std::vector<TreeNode> &finestOctree, unsigned dof = 1, double loadFac = 1.5,
                              bool compressLut = true, bool incCorner = true) {
    (void)loadFac;
    (void)compressLut;
    (void)incCorner;
    dgx::require_one_rank(comm);
    if (dof != 1) throw dgx::Error("DAMGCreateAndSetDA: only dof == 1 is supported");
    dgx::ensure_init();
    const unsigned md = finestOctree.empty() ? 0 : finestOctree[0].getMaxDepth();
    dgx::detail::Octree o = dgx::detail::upload(finestOctree, md, false);
    dgx_mg mg = nullptr;
    dgx::check(dgx_mg_create(o.get(), nlevels, 1.0, 1.0, &mg), "DAMGCreateAndSetDA");
    int nl = 0;
    dgx::check(dgx_mg_info(mg, &nl, nullptr, nullptr), "DAMGCreateAndSetDA");
    nlevels = nl;
    DAMG *arr = new DAMG[nl];
    for (int i = 0; i < nl; ++i) {
        arr[i] = new _p_DAMG();
        arr[i]->dof = dof;
        arr[i]->nlevels = nl - i;
        arr[i]->totalLevels = nl;
        arr[i]->comm = comm;
        arr[i]->user = user;
        arr[i]->mg = mg;
        arr[i]->level = nl - 1 - i;
    }
    dgx_da fine = nullptr;
    dgx::check(dgx_mg_finest_da(mg, &fine), "DAMGCreateAndSetDA");
    arr[nl - 1]->da = new DA(fine);
    arr[nl - 1]->x.assign(arr[nl - 1]->da->getNodeSize(), 0.0);
    arr[nl - 1]->b.assign(arr[nl - 1]->da->getNodeSize(), 0.0);
    finestOctree.clear();
    *damg = arr;
    return 0;
}

// DAMGCreateSuppressedDOFs (omg.h:504; blockDiag.cpp:46-66): allocates each
// level's suppressedDOF, dof * getLocalBufferSize() bytes.  The reference
// leaves filling it to the Dirichlet driver (SetDirichletJacContexts,
// omgJac.C:509-587: assignBoundaryFlags copied to the node indices of every
// element); here it is filled with the same flags at once (FACE_BDY 1,
// EDGE_BDY 2, CORNER_BDY 3 for boundary nodes, 0 elsewhere), since the device
// operator already applies them.  Returns 1 like the reference.
inline int DAMGCreateSuppressedDOFs(DAMG *damg) {
    const int nl = damg[0]->nlevels;
    for (int i = 0; i < nl; ++i) {
        dgx_da d = nullptr;
        dgx::check(dgx_mg_level_da(damg[i]->mg, damg[i]->level, &d), "DAMGCreateSuppressedDOFs");
        dgx_da_info inf{};
        dgx::check(dgx_da_get_info(d, &inf), "DAMGCreateSuppressedDOFs");
        const size_t sz = (size_t)damg[i]->dof * inf.local_buffer_size;
        delete[] damg[i]->suppressedDOF;
        damg[i]->suppressedDOF = nullptr;
        if (!sz) continue;
        damg[i]->suppressedDOF = new unsigned char[sz];
        dgx::check(dgx_da_boundary_flags(d, damg[i]->suppressedDOF), "DAMGCreateSuppressedDOFs");
    }
    return 1;
}

// DAMGSetKSP (omg.h:488): operator -lapFac lap + massFac I with homogeneous
// Dirichlet rows (omgJac.C:958-1057) and the right-hand-side callback, which
// fills b (getNodeSize() entries, the load vector including the mass matrix)
// on the finest level.
inline int DAMGSetKSP(DAMG *damg, double lapFac, double massFac, int (*rhs)(DAMG, std::vector<double> &)) {
    const int nl = damg[0]->nlevels;
    for (int i = 0; i < nl; ++i) {
        damg[i]->lapFac = lapFac;
        damg[i]->massFac = massFac;
        damg[i]->rhs = rhs;
    }
    dgx::check(dgx_mg_set_operator(damg[0]->mg, lapFac, massFac), "DAMGSetKSP");
    return 0;
}

// Load vector b = D M f of a nodal f (zero on Dirichlet nodes), for rhs
// callbacks (the reference drivers assemble it with the mass matrix).
inline void DAMGLoadVector(DAMG damg, const std::vector<double> &f, std::vector<double> &b) {
    const size_t n = f.size(), bytes = n * sizeof(double);
    void *df = nullptr, *db = nullptr;
    dgx::check(dgx_device_malloc(&df, bytes ? bytes : 8), "malloc");
    dgx::check(dgx_device_malloc(&db, bytes ? bytes : 8), "malloc");
    int rc = dgx_memcpy_htod(df, f.data(), bytes);
    if (!rc) rc = dgx_mg_rhs(damg->mg, (const double *)df, (double *)db);
    b.resize(n);
    if (!rc) rc = dgx_memcpy_dtoh(b.data(), db, bytes);
    dgx_device_free(df);
    dgx_device_free(db);
    dgx::check(rc, "DAMGLoadVector");
}

// DAMGSolve (omg.h:492; blockDiag.cpp:1122-1170): CG on the finest level,
// preconditioned by one V-cycle per iteration (FP32 V-cycle when
// opts.fp32_vcycle, the default), to ||b - A x|| <= opts.rtol ||b||.
inline int DAMGSolve(DAMG *damg) {
    DAMG fine = damg[damg[0]->nlevels - 1];
    if (fine->rhs) dgx::check(fine->rhs(fine, fine->b), "DAMGSolve rhs");
    const size_t n = fine->da->getNodeSize(), bytes = n * sizeof(double);
    if (fine->b.size() != n) throw dgx::Error("DAMGSolve: b must have getNodeSize() entries");
    void *db = nullptr, *dx = nullptr;
    dgx::check(dgx_device_malloc(&db, bytes ? bytes : 8), "malloc");
    dgx::check(dgx_device_malloc(&dx, bytes ? bytes : 8), "malloc");
    int rc = dgx_memcpy_htod(db, fine->b.data(), bytes);
    if (!rc) rc = dgx_mg_solve(fine->mg, (const double *)db, (double *)dx, &fine->opts, &fine->stats);
    fine->x.resize(n);
    if (!rc) rc = dgx_memcpy_dtoh(fine->x.data(), dx, bytes);
    dgx_device_free(db);
    dgx_device_free(dx);
    dgx::check(rc, "DAMGSolve");
    return 0;
}

inline int DAMGDestroy(DAMG *damg) {
    if (!damg) return 0;
    const int nl = damg[0]->nlevels;
    dgx_mg mg = damg[0]->mg;
    for (int i = 0; i < nl; ++i) {
        delete damg[i]->da;
        delete[] damg[i]->suppressedDOF;
        delete[] damg[i]->suppressedDOFaux;
        delete damg[i];
    }
    delete[] damg;
    if (mg) dgx::check(dgx_mg_free(mg), "DAMGDestroy");
    return 0;
}

}  // namespace ot

// Accessors (omg.h:531-606).
#define DAMGGetRHS(ctx) (ctx)[(ctx)[0]->nlevels - 1]->b
#define DAMGGetx(ctx) (ctx)[(ctx)[0]->nlevels - 1]->x
#define DAMGGetComm(ctx) (ctx)[(ctx)[0]->nlevels - 1]->comm
#define DAMGGetFine(ctx) (ctx)[(ctx)[0]->nlevels - 1]
#define DAMGGetDA(ctx) ((ctx)[(ctx)[0]->nlevels - 1]->da)
#define DAMGGetUser(ctx, level) ((ctx)[level]->user)
#define DAMGGetLevels(ctx) (ctx)[0]->nlevels
#define DAMGGetDAMG(ctx) (ctx)[(ctx)[0]->nlevels - 1]

// treeNodesTovtk (treenode2vtk.h:19, treenode2vtk.cpp:4-134): sorts nodes
// along the curve and writes <name>_H_<rank>.vtk (_M_ for Morton).
inline void treeNodesTovtk(std::vector<ot::TreeNode> &nodes, int mpi_rank, std::string vtk_file_name,
                           bool hsorted = false) {
    dgx::check(dgx_write_vtk(reinterpret_cast<uint32_t *>(nodes.data()), nodes.size(), mpi_rank,
                             vtk_file_name.c_str(), hsorted ? 1 : 0, dgx::curve()),
               "treeNodesTovtk");
}

namespace par {
// par::sampleSort (parUtils.h:424-425) on one rank: a device sort of the
// TreeNodes along the curve.
template <typename T>
int sampleSort(std::vector<T> &in, std::vector<T> &out, MPI_Comm comm) {
    static_assert(std::is_same<T, ot::TreeNode>::value, "par::sampleSort is provided for ot::TreeNode");
    dgx::require_one_rank(comm);
    dgx::ensure_init();
    const unsigned md = in.empty() ? 0 : in[0].getMaxDepth();
    dgx::detail::Octree o = dgx::detail::upload(in, md, true);
    dgx::detail::download(o, out);
    return 1;
}
}  // namespace par

namespace dgx {
namespace detail {
inline Octree upload(const std::vector<ot::TreeNode> &v, unsigned maxDepth, bool sort) {
    Octree o;
    check(dgx_octree_import(reinterpret_cast<const uint32_t *>(v.data()), v.size(), maxDepth, curve(), 0,
                            sort ? 1 : 0, o.out()),
          "import");
    return o;
}
inline void download(const Octree &o, std::vector<ot::TreeNode> &v) {
    size_t n = 0;
    check(dgx_octree_size(o.get(), &n), "size");
    v.resize(n);
    check(dgx_octree_export(o.get(), reinterpret_cast<uint32_t *>(v.data()), 0), "export");
}
}  // namespace detail
}  // namespace dgx
