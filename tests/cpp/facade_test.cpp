// GPU test of the C++ facade (include/dendro_mi355x.hpp), written the way a
// reference driver (tstMatVec.C / tstDirichlet.C) calls ot:: / par:: / DAMG.
// The oracle (CPU restatement, test infrastructure) is the checker.
// Prints "FACADE OK" and exits 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/dendro_mi355x.hpp"
#include "../../oracle/dendro_oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                 \
    do {                                                 \
        if (!(cond)) {                                   \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                    \
            std::printf("\n");                           \
            ++failures;                                  \
        }                                                \
    } while (0)

static bool same_octants(const std::vector<ot::TreeNode> &a, const or_tn *b, size_t nb) {
    if (a.size() != nb) return false;
    for (size_t i = 0; i < nb; ++i)
        if (a[i].m_uiX != b[i].x || a[i].m_uiY != b[i].y || a[i].m_uiZ != b[i].z ||
            a[i].getLevel() != (b[i].lev & 31u))
            return false;
    return true;
}

static std::vector<double> g_f;  // nodal right-hand side for the DAMG callback

static int rhs_callback(ot::DAMG damg, std::vector<double> &b) {
    ot::DAMGLoadVector(damg, g_f, b);
    return 0;
}

static int run() {
    ot::_InitializeHcurve(3);
    const unsigned maxDepth = 30, maxNumPts = 8;
    const size_t n = 30000;
    std::vector<double> pts(3 * n);
    dgx::check(dgx_gen_points(1, 5, n, 0, 0, 1, 0, 0, 0, pts.data()), "gen");
    std::vector<double> pts_copy = pts;

    // points2Octree
    double gLens[3] = {1.0, 1.0, 1.0};
    std::vector<ot::TreeNode> linOct;
    ot::points2Octree(pts, gLens, linOct, 3, maxDepth, maxNumPts, MPI_COMM_WORLD);
    CHECK(pts.empty(), "points2Octree must clear pts");
    or_tn *ro = nullptr;
    size_t nro = 0;
    CHECK(or_points2octree(pts_copy.data(), n, maxDepth, maxNumPts, &ro, &nro) == 0, "oracle p2o");
    CHECK(same_octants(linOct, ro, nro), "points2Octree differs from the oracle (%zu vs %zu)", linOct.size(), nro);
    CHECK(std::is_sorted(linOct.begin(), linOct.end()), "TreeNode::operator< disagrees with the output order");

    // par::sampleSort of a shuffled copy
    std::vector<ot::TreeNode> shuf = linOct, sorted;
    std::mt19937 rng(7);
    std::shuffle(shuf.begin(), shuf.end(), rng);
    par::sampleSort(shuf, sorted, MPI_COMM_WORLD);
    CHECK(sorted == linOct, "sampleSort does not restore the order");

    // balanceOctree
    std::vector<ot::TreeNode> balOct, linCopy = linOct;
    ot::balanceOctree(linOct, balOct, 3, maxDepth, true, MPI_COMM_WORLD);
    CHECK(linOct.empty(), "balanceOctree must clear its input");
    or_tn *rb = nullptr;
    size_t nrb = 0;
    CHECK(or_balance(ro, nro, 1, 0, &rb, &nrb) == 0, "oracle balance");
    CHECK(same_octants(balOct, rb, nrb), "balanceOctree differs from the oracle (%zu vs %zu)", balOct.size(), nrb);

    // DA: sizes, maps through the iteration interface, child numbers
    std::vector<ot::TreeNode> daIn = balOct;
    ot::DA da(daIn, MPI_COMM_WORLD, MPI_COMM_WORLD, 0.3, false);
    CHECK(daIn.empty(), "DA must consume its input");
    or_da rda;
    CHECK(or_da_build(rb, nrb, &rda) == 0, "oracle DA");
    CHECK(da.getElementSize() == rda.n_elem, "element size %u vs %zu", da.getElementSize(), rda.n_elem);
    CHECK(da.getLocalBufferSize() == rda.n_local, "local buffer size");
    CHECK(da.getNodeSize() == rda.n_nodes, "node size %u vs %u", da.getNodeSize(), rda.n_nodes);
    CHECK(da.getBoundaryNodeSize() == rda.bnd_node_size, "boundary node size");
    size_t bad_nlist = 0, bad_mask = 0, bad_child = 0, elems = 0;
    for (da.init<ot::DA_FLAGS::ALL>(); da.curr() < da.end<ot::DA_FLAGS::ALL>(); da.next<ot::DA_FLAGS::ALL>()) {
        const unsigned e = da.curr();
        unsigned idx[8];
        da.getNodeIndices(idx);
        for (int j = 0; j < 8; ++j)
            if (idx[j] != rda.nlist[8 * (size_t)e + j]) ++bad_nlist;
        if (da.getHangingNodeIndex(e) != rda.mask[e]) ++bad_mask;
        // child number from the oracle octant: Morton child at its level
        const or_tn &o = rda.octs[e];
        const unsigned lev = o.lev & 31u, b = da.getMaxDepth() - lev;
        const unsigned c = ((o.x >> b) & 1u) | (((o.y >> b) & 1u) << 1) | (((o.z >> b) & 1u) << 2);
        if (da.getChildNumber() != c) ++bad_child;
        ++elems;
    }
    CHECK(elems == rda.n_elem, "iteration visited %zu elements", elems);
    CHECK(bad_nlist == 0, "%zu nlist entries differ", bad_nlist);
    CHECK(bad_mask == 0, "%zu hanging masks differ", bad_mask);
    CHECK(bad_child == 0, "%zu child numbers differ", bad_child);
    size_t bad_flags = 0;
    for (unsigned i = 0; i < da.getLocalBufferSize(); ++i)
        if (da.getFlag(i) != (rda.octs[i].lev & 0xffu)) ++bad_flags;
    CHECK(bad_flags == 0, "%zu octant flags differ", bad_flags);

    // matvec invariants of -lap + I on the free (Neumann) operator: the
    // Laplacian rows sum to 0 and the mass matrix sums to the domain volume 1.
    const unsigned nn = da.getNodeSize();
    std::vector<double> one(nn, 1.0), yk, ym;
    da.matvec(one, yk, 1.0, 1.0);
    da.matvec(one, ym, 0.0, 1.0);
    double dmax = 0.0, msum = 0.0, mmax = 0.0;
    for (unsigned i = 0; i < nn; ++i) {
        dmax = std::max(dmax, std::fabs(yk[i] - ym[i]));
        mmax = std::max(mmax, std::fabs(ym[i]));
        msum += ym[i];
    }
    CHECK(dmax <= 1e-12 * mmax, "K 1 != 0: %g", dmax);
    CHECK(std::fabs(msum - 1.0) <= 1e-12, "sum(M 1) = %.17g", msum);
    // ghosted view round trip
    std::vector<double> v(nn);
    for (unsigned i = 0; i < nn; ++i) v[i] = i;
    double *buf = nullptr;
    da.vecGetBuffer(v, buf, false, false, true);
    const auto &l2n = da.localToNode();
    size_t bad_view = 0;
    for (unsigned i = 0; i < da.getLocalBufferSize(); ++i)
        if (l2n[i] != 0xffffffffu && buf[i] != (double)l2n[i]) ++bad_view;
    da.vecRestoreBuffer(v, buf, false, false, true);
    CHECK(bad_view == 0, "vecGetBuffer view differs at %zu entries", bad_view);

    // DAMG: Dirichlet Poisson solve through the reference's entry points
    std::vector<ot::TreeNode> finest = balOct;
    int nlevels = 3;
    ot::DAMG *damg = nullptr;
    ot::DAMG_Initialize(MPI_COMM_WORLD);
    ot::DAMGCreateAndSetDA(MPI_COMM_WORLD, nlevels, nullptr, &damg, finest);
    CHECK(finest.empty(), "DAMGCreateAndSetDA must clear the finest octree");
    CHECK(nlevels >= 2, "levels built: %d", nlevels);
    ot::DA *fda = DAMGGetDA(damg);
    // suppressed DOFs: on every level, the count of boundary axes of each
    // node's position (assignBoundaryFlags), 0 for entries that are not nodes
    CHECK(ot::DAMGCreateSuppressedDOFs(damg) == 1, "DAMGCreateSuppressedDOFs");
    for (int lv = 0; lv < nlevels; ++lv)
        CHECK(damg[lv]->suppressedDOF != nullptr && damg[lv]->suppressedDOFaux == nullptr, "level %d arrays", lv);
    {
        // against the oracle's DA (same local buffer): its octants' anchors
        // in the embedded domain [0, 2^(maxDepth+1)), the unit cube ends at 2^maxDepth
        const ot::DAMG fl = DAMGGetFine(damg);
        CHECK(fda->getLocalBufferSize() == rda.n_local, "local buffer sizes differ");
        size_t bad_sup = 0, marked = 0;
        for (size_t i = 0; i < rda.n_local; ++i) {
            const or_tn &o = rda.octs[i];
            const unsigned top = 1u << (o.md - 1);
            unsigned want = 0;
            if (rda.local_to_node[i] != 0xffffffffu)
                want = (o.x == 0 || o.x == top) + (o.y == 0 || o.y == top) + (o.z == 0 || o.z == top);
            bad_sup += fl->suppressedDOF[i] != want;
            marked += want != 0;
        }
        CHECK(bad_sup == 0 && marked > 0, "suppressedDOF differs at %zu of %u entries (%zu marked)", bad_sup,
              fda->getLocalBufferSize(), marked);
    }
    g_f.assign(fda->getNodeSize(), 0.0);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (auto &x : g_f) x = nd(rng);
    ot::DAMGSetKSP(damg, 1.0, 1.0, rhs_callback);
    ot::DAMGSolve(damg);
    const ot::DAMG fine = DAMGGetFine(damg);
    const int cg_iterations = fine->stats.iterations;
    CHECK(fine->stats.rel_residual <= 1e-8, "relative residual %.3e", fine->stats.rel_residual);
    CHECK(fine->stats.iterations >= 1 && fine->stats.iterations < 50, "iterations %d", fine->stats.iterations);
    // independent residual through the Dirichlet operator of the finest DA
    const size_t fn = fda->getNodeSize(), bytes = fn * sizeof(double);
    void *dx = nullptr, *dy = nullptr;
    dgx::check(dgx_device_malloc(&dx, bytes), "malloc");
    dgx::check(dgx_device_malloc(&dy, bytes), "malloc");
    dgx::check(dgx_memcpy_htod(dx, DAMGGetx(damg).data(), bytes), "copy");
    dgx::check(dgx_da_matvec_dirichlet(fda->handle(), (const double *)dx, (double *)dy, 1.0, 1.0), "matvec");
    std::vector<double> ax(fn);
    dgx::check(dgx_memcpy_dtoh(ax.data(), dy, bytes), "copy");
    dgx_device_free(dx);
    dgx_device_free(dy);
    double rr = 0.0, bb = 0.0;
    for (size_t i = 0; i < fn; ++i) {
        const double r = DAMGGetRHS(damg)[i] - ax[i];
        rr += r * r;
        bb += DAMGGetRHS(damg)[i] * DAMGGetRHS(damg)[i];
    }
    CHECK(std::sqrt(rr / bb) <= 1.5e-8, "recomputed residual %.3e", std::sqrt(rr / bb));
    ot::DAMGDestroy(damg);
    ot::DAMG_Finalize();

    // compressLut with Morton order: getNodeIndices decodes the reference's
    // compressed maps (host) to the plain node list of the same octree
    {
        dgx::curve() = DGX_CURVE_MORTON;
        std::vector<double> mp(pts_copy);
        std::vector<ot::TreeNode> mo, mb;
        double gl[3] = {1.0, 1.0, 1.0};
        ot::points2Octree(mp, gl, mo, 3, maxDepth, maxNumPts, MPI_COMM_WORLD);
        ot::balanceOctree(mo, mb, 3, maxDepth, true, MPI_COMM_WORLD);
        std::vector<ot::TreeNode> mb2 = mb;
        ot::DA plain(mb, MPI_COMM_WORLD, MPI_COMM_WORLD, 0.1, false);
        ot::DA packed(mb2, MPI_COMM_WORLD, MPI_COMM_WORLD, 0.1, true);
        CHECK(packed.isLutCompressed() && !plain.isLutCompressed(), "compressLut not honoured");
        size_t bad_lut = 0, elems_lut = 0;
        for (packed.init<ot::DA_FLAGS::ALL>(), plain.init<ot::DA_FLAGS::ALL>();
             packed.curr() < packed.end<ot::DA_FLAGS::ALL>(); packed.next<ot::DA_FLAGS::ALL>(),
                                                          plain.next<ot::DA_FLAGS::ALL>()) {
            unsigned a[8], b[8];
            packed.getNodeIndices(a);
            plain.getNodeIndices(b);
            bad_lut += !std::equal(a, a + 8, b);
            ++elems_lut;
        }
        CHECK(elems_lut == plain.getElementSize() && bad_lut == 0, "compressed getNodeIndices differ at %zu of %zu",
              bad_lut, elems_lut);
        dgx::curve() = DGX_CURVE_HILBERT;
    }

    or_da_free(&rda);
    or_free(ro);
    or_free(rb);
    if (failures) {
        std::printf("FACADE FAILED (%d checks)\n", failures);
        return 1;
    }
    std::printf("FACADE OK: %zu octants, %zu balanced, %u nodes, %d levels, %d CG iterations\n", linCopy.size(),
                balOct.size(), nn, nlevels, cg_iterations);
    return 0;
}

int main() {
    try {
        return run();
    } catch (const std::exception &e) {
        std::printf("FACADE ERROR: %s\n", e.what());
        return 2;
    }
}
