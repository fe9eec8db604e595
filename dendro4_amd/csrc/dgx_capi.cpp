// dgx_capi.cpp -- the extern "C" boundary (include/dendro_mi355x.h): runtime
// context, error reporting, handles and argument checks.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <csignal>
#include <cstdlib>
#include <execinfo.h>
#include <unistd.h>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>
#include <string>

#include "../../include/dendro_mi355x.h"
#include "dgx_mg.h"

// ------------------------------------------------------------------ state --
namespace {
thread_local std::string g_err;
struct Ctx {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipMemPool_t pool = nullptr;
};
Ctx g_ctx;
std::mutex g_init_mu;
}  // namespace

void dgx_set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}


// Device of the initialised context (internal, for the request server).
bool dgx_ctx_device(int *device) {
    if (!g_ctx.ready) return false;
    *device = g_ctx.device;
    return true;
}

static void ensure_ctx() {
    if (!g_ctx.ready) {
        dgx_set_error("dgx_init() has not been called");
        throw DgxError();
    }
}
static hipStream_t S() { return g_ctx.stream; }

#define API_BEGIN try {
#define API_END                                                         \
    sort_fault_check();                                                 \
    }                                                                   \
    catch (DgxError &) {                                                \
        return -1;                                                      \
    }                                                                   \
    catch (std::exception & ex) {                                       \
        dgx_set_error("%s", ex.what());                                 \
        return -1;                                                      \
    }                                                                   \
    return 0;

struct dgx_octree_s {
    DevBuf<Key> keys;
    size_t n = 0;
    uint32_t md = 30;
    int curve = CURVE_HILBERT;
    uint64_t stats[5] = {0, 0, 0, 0, 0};
};
struct dgx_da_s {
    std::unique_ptr<DaDevice> da;
    DaDevice *borrowed = nullptr;  // a DA owned by a multigrid hierarchy
    const DaDevice &get() const { return borrowed ? *borrowed : *da; }
};
struct dgx_mg_s {
    std::unique_ptr<MgHierarchy> mg;
    dgx_da_s finest;
    std::vector<std::unique_ptr<dgx_da_s>> levels;  // borrowed views of every level (dgx_mg_level_da)
};

extern "C" {

// With DGX_SEGV_BACKTRACE set, a fatal signal prints the native stack to
// stderr before the default action (diagnostics for host-side crashes).
static void segv_backtrace(int sig) {
    void *fr[64];
    const int n = backtrace(fr, 64);
    const char msg[] = "dendro_mi355x: fatal signal, native backtrace:\n";
    (void)!write(2, msg, sizeof msg - 1);
    backtrace_symbols_fd(fr, n, 2);
    signal(sig, SIG_DFL);
    raise(sig);
}

int dgx_init(int device) {
    if (std::getenv("DGX_SEGV_BACKTRACE")) {
        signal(SIGSEGV, segv_backtrace);
        signal(SIGABRT, segv_backtrace);
    }
    std::lock_guard<std::mutex> lk(g_init_mu);
    API_BEGIN
    if (g_ctx.ready) {
        if (g_ctx.device != device) {
            dgx_set_error("dgx_init: already initialised on device %d", g_ctx.device);
            return -1;
        }
        return 0;
    }
    int ndev = 0;
    HIP_OK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        dgx_set_error("dgx_init: device %d not present (%d devices)", device, ndev);
        return -1;
    }
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    HIP_OK(hipDeviceGetDefaultMemPool(&g_ctx.pool, device));
    // The pool keeps up to an eighth of the device of freed memory mapped (as
    // the caching allocator keeps up to an eighth in blocks); beyond that it
    // returns memory to the device at synchronisation points, so the runtime's
    // own allocations (kernel scratch, library workspaces) still find room
    // beside a request server's budget of three quarters.  Memory returned to
    // the device has to be mapped again by the next allocation that needs it
    // (several milliseconds per GB): with 4 GiB here, the request server with
    // resident problems, whose evictions free whole hierarchies, spent 16 s of
    // its 2,000-request stream in hipMallocAsync.
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    uint64_t thr = (uint64_t)tot / 8;
    HIP_OK(hipMemPoolSetAttribute(g_ctx.pool, hipMemPoolAttrReleaseThreshold, &thr));
    // device address of the static scratch array of dgx_da_matvec_noalloc
    da_noalloc_init();
    g_ctx.device = device;
    g_ctx.ready = true;
    API_END
}

const char *dgx_last_error(void) { return g_err.c_str(); }

void *dgx_stream(void) { return g_ctx.ready ? (void *)g_ctx.stream : nullptr; }

int dgx_synchronize(void) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}

int dgx_memory_info(uint64_t out[5], int reset_peak) {
    API_BEGIN
    ensure_ctx();
    if (!out) { dgx_set_error("memory info: null argument"); return -1; }
    size_t fr = 0, tot = 0, live = 0, peak = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    dgx_alloc_live(&live, &peak, reset_peak);
    out[0] = tot;
    out[1] = tot - fr;
    out[2] = live;
    out[3] = peak;
    out[4] = dgx_alloc_cached_bytes();
    API_END
}

int dgx_trim_cache(void) {
    API_BEGIN
    ensure_ctx();
    dgx_alloc_trim();
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemPoolTrimTo(g_ctx.pool, 0));
    API_END
}

int dgx_device_malloc(void **p, size_t bytes) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipMalloc(p, bytes));
    API_END
}
int dgx_device_free(void *p) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipStreamSynchronize(S()));
    HIP_OK(hipFree(p));
    API_END
}
int dgx_memcpy_htod(void *dst, const void *src, size_t bytes) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S()));
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}
int dgx_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S()));
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}
int dgx_memset_device(void *dst, int value, size_t bytes) {
    API_BEGIN
    ensure_ctx();
    HIP_OK(hipMemsetAsync(dst, value, bytes, S()));
    API_END
}

int dgx_points_to_octree_glens(const double *xyz, size_t n, const double *glens, unsigned max_depth,
                               unsigned max_num_pts, int curve, int on_device, dgx_octree *out) {
    API_BEGIN
    ensure_ctx();
    if (!out || (n && !xyz)) { dgx_set_error("points2Octree: null argument"); return -1; }
    if (max_depth > 30) { dgx_set_error("points2Octree: maxDepth %u > 30", max_depth); return -1; }
    if (max_num_pts == 0) { dgx_set_error("points2Octree: maxNumPts must be >= 1"); return -1; }
    if (curve != DGX_CURVE_HILBERT && curve != DGX_CURVE_MORTON) { dgx_set_error("unknown curve %d", curve); return -1; }
    if (n > 0xffffffffull) { dgx_set_error("points2Octree: %zu points exceed 2^32-1", n); return -1; }
    if (glens && !(glens[0] > 0.0 && glens[1] > 0.0 && glens[2] > 0.0)) {
        dgx_set_error("points2Octree: domain lengths must be positive");
        return -1;
    }
    std::unique_ptr<dgx_octree_s> o(new dgx_octree_s());
    o->md = max_depth;
    o->curve = curve;
    DevBuf<double> tmp;
    const double *d = xyz;
    if (!on_device && n) {
        tmp.alloc(3 * n, S());
        HIP_OK(hipMemcpyAsync(tmp.p, xyz, 3 * n * sizeof(double), hipMemcpyHostToDevice, S()));
        d = tmp.p;
    }
    P2OStats st;
    o->keys = p2o_build(d, n, max_depth, max_num_pts, curve, S(), &st, glens);
    o->n = o->keys.n;
    o->stats[0] = st.textbook_leaves;
    o->stats[1] = st.triggers;
    o->stats[2] = st.replays;
    o->stats[3] = st.replayed_leaves;
    *out = o.release();
    API_END
}

int dgx_points_to_octree(const double *xyz, size_t n, unsigned max_depth, unsigned max_num_pts, int curve,
                         int on_device, dgx_octree *out) {
    return dgx_points_to_octree_glens(xyz, n, nullptr, max_depth, max_num_pts, curve, on_device, out);
}

int dgx_balance_octree(dgx_octree in, int inc_corner, dgx_octree *out) {
    (void)inc_corner;
    API_BEGIN
    ensure_ctx();
    if (!in || !out) { dgx_set_error("balanceOctree: null argument"); return -1; }
    std::unique_ptr<dgx_octree_s> o(new dgx_octree_s());
    o->md = in->md;
    o->curve = in->curve;
    BalStats st;
    o->keys = balance_build(in->keys.p, in->n, in->md, in->curve, S(), &st);
    o->n = o->keys.n;
    o->stats[0] = st.levels;
    o->stats[1] = st.seeds;
    *out = o.release();
    API_END
}

int dgx_octree_import(const uint32_t *tn, size_t n, unsigned max_depth, int curve, int on_device, int sort,
                      dgx_octree *out) {
    API_BEGIN
    ensure_ctx();
    if (!out || (n && !tn)) { dgx_set_error("import: null argument"); return -1; }
    if (max_depth > 31) { dgx_set_error("import: maxDepth %u > 31", max_depth); return -1; }
    if (curve != DGX_CURVE_HILBERT && curve != DGX_CURVE_MORTON) { dgx_set_error("unknown curve %d", curve); return -1; }
    std::unique_ptr<dgx_octree_s> o(new dgx_octree_s());
    o->md = max_depth;
    o->curve = curve;
    DevBuf<uint32_t> tmp;
    const uint32_t *d = tn;
    if (!on_device && n) {
        tmp.alloc(7 * n, S());
        HIP_OK(hipMemcpyAsync(tmp.p, tn, 7 * n * sizeof(uint32_t), hipMemcpyHostToDevice, S()));
        d = tmp.p;
    }
    o->keys = import_treenodes(d, n, max_depth, curve, sort != 0, S());
    o->n = n;
    *out = o.release();
    API_END
}

int dgx_octree_size(dgx_octree o, size_t *n) {
    if (!o || !n) { dgx_set_error("null argument"); return -1; }
    *n = o->n;
    return 0;
}
int dgx_octree_max_depth(dgx_octree o, unsigned *md) {
    if (!o || !md) { dgx_set_error("null argument"); return -1; }
    *md = o->md;
    return 0;
}
int dgx_octree_stats(dgx_octree o, uint64_t out[5]) {
    if (!o || !out) { dgx_set_error("null argument"); return -1; }
    memcpy(out, o->stats, sizeof(o->stats));
    return 0;
}

int dgx_octree_export(dgx_octree o, uint32_t *tn, int on_device) {
    API_BEGIN
    ensure_ctx();
    if (!o || (o->n && !tn)) { dgx_set_error("export: null argument"); return -1; }
    if (!o->n) return 0;
    if (on_device) {
        export_treenodes(o->keys.p, o->n, o->md, o->curve, tn, S());
        HIP_OK(hipStreamSynchronize(S()));
    } else {
        DevBuf<uint32_t> tmp(7 * o->n, S());
        export_treenodes(o->keys.p, o->n, o->md, o->curve, tmp.p, S());
        HIP_OK(hipMemcpyAsync(tn, tmp.p, 7 * o->n * sizeof(uint32_t), hipMemcpyDeviceToHost, S()));
        HIP_OK(hipStreamSynchronize(S()));
    }
    API_END
}

int dgx_octree_free(dgx_octree o) {
    API_BEGIN
    if (o) {
        if (g_ctx.ready) HIP_OK(hipStreamSynchronize(S()));
        delete o;
    }
    API_END
}

int dgx_da_create(dgx_octree oct, dgx_da *out) {
    API_BEGIN
    ensure_ctx();
    if (!oct || !out) { dgx_set_error("DA: null argument"); return -1; }
    if (oct->md > 30) { dgx_set_error("DA: maxDepth %u > 30 cannot be embedded", oct->md); return -1; }
    std::unique_ptr<dgx_da_s> d(new dgx_da_s());
    d->da = da_build(oct->keys.p, oct->n, oct->md, oct->curve, S());
    *out = d.release();
    API_END
}

int dgx_da_get_info(dgx_da d, dgx_da_info *info) {
    if (!d || !info) { dgx_set_error("null argument"); return -1; }
    const DaDevice &a = d->get();
    info->local_buffer_size = a.n_local;
    info->element_size = a.n_elem;
    info->node_size = a.node_size;
    info->boundary_node_size = a.bnd_node_size;
    info->n_nodes = a.n_nodes;
    info->max_depth = a.md;
    info->ghost_scans = a.n_ghost_scans;
    info->matvec_blocks = a.n_blocks;
    info->block_node_pairs = a.n_pairs;
    info->shared_nodes = a.n_shared;
    info->shared_pairs = a.partial_slots;
    info->max_block_nodes = a.max_block_nodes;
    return 0;
}

int dgx_da_export(dgx_da d, uint32_t *nlist, uint8_t *masks, uint8_t *oct_levels, uint32_t *l2n) {
    API_BEGIN
    ensure_ctx();
    if (!d) { dgx_set_error("null argument"); return -1; }
    const DaDevice &a = d->get();
    if (nlist) HIP_OK(hipMemcpyAsync(nlist, a.nlist.p, 8 * a.n_elem * 4, hipMemcpyDeviceToHost, S()));
    if (masks) HIP_OK(hipMemcpyAsync(masks, a.mask.p, a.n_elem, hipMemcpyDeviceToHost, S()));
    if (oct_levels) HIP_OK(hipMemcpyAsync(oct_levels, a.flags.p, a.n_local, hipMemcpyDeviceToHost, S()));
    if (l2n) HIP_OK(hipMemcpyAsync(l2n, a.l2n.p, a.n_local * 4, hipMemcpyDeviceToHost, S()));
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}

int dgx_da_export_anchors(dgx_da d, uint32_t *xyz) {
    API_BEGIN
    ensure_ctx();
    if (!d || !xyz) { dgx_set_error("null argument"); return -1; }
    const DaDevice &a = d->get();
    std::vector<uint32_t> t(a.n_local);
    const uint32_t *src[3] = {a.x.p, a.y.p, a.z.p};
    for (int c = 0; c < 3; ++c) {
        if (a.n_local) HIP_OK(hipMemcpyAsync(t.data(), src[c], a.n_local * 4, hipMemcpyDeviceToHost, S()));
        HIP_OK(hipStreamSynchronize(S()));
        for (size_t i = 0; i < a.n_local; ++i) xyz[3 * i + c] = t[i];
    }
    API_END
}

int dgx_da_boundary_flags(dgx_da d, uint8_t *flags) {
    API_BEGIN
    ensure_ctx();
    if (!d || !flags) { dgx_set_error("null argument"); return -1; }
    const DaDevice &a = d->get();
    std::vector<uint32_t> x(a.n_local), y(a.n_local), z(a.n_local), l2n(a.n_local);
    if (a.n_local) {
        HIP_OK(hipMemcpyAsync(x.data(), a.x.p, a.n_local * 4, hipMemcpyDeviceToHost, S()));
        HIP_OK(hipMemcpyAsync(y.data(), a.y.p, a.n_local * 4, hipMemcpyDeviceToHost, S()));
        HIP_OK(hipMemcpyAsync(z.data(), a.z.p, a.n_local * 4, hipMemcpyDeviceToHost, S()));
        HIP_OK(hipMemcpyAsync(l2n.data(), a.l2n.p, a.n_local * 4, hipMemcpyDeviceToHost, S()));
    }
    HIP_OK(hipStreamSynchronize(S()));
    // the domain is [0, 2^(md-1)] in the embedded coordinates
    const uint32_t top = 1u << (a.md - 1);
    for (size_t i = 0; i < a.n_local; ++i) {
        if (l2n[i] == 0xffffffffu) {
            flags[i] = 0;
            continue;
        }
        flags[i] = (uint8_t)((x[i] == 0 || x[i] == top) + (y[i] == 0 || y[i] == top) + (z[i] == 0 || z[i] == top));
    }
    API_END
}

int dgx_da_compress_maps(dgx_da d, uint8_t *rem, uint8_t *qmask, uint8_t *order, uint16_t *quot, size_t quot_cap,
                         size_t *n_quot) {
    API_BEGIN
    ensure_ctx();
    if (!d || !n_quot) { dgx_set_error("null argument"); return -1; }
    const DaDevice &a = d->get();
    LutMaps m;
    const size_t nq = da_lut_compress(a, m, S());
    *n_quot = nq;
    if (quot && quot_cap < nq) {
        dgx_set_error("compressed maps: %zu quotients do not fit a buffer of %zu", nq, quot_cap);
        return -1;
    }
    const size_t ne = a.n_elem;
    if (rem && ne) HIP_OK(hipMemcpyAsync(rem, m.rem.p, 8 * ne, hipMemcpyDeviceToHost, S()));
    if (qmask && ne) HIP_OK(hipMemcpyAsync(qmask, m.qmask.p, ne, hipMemcpyDeviceToHost, S()));
    if (order && ne) HIP_OK(hipMemcpyAsync(order, m.order.p, ne, hipMemcpyDeviceToHost, S()));
    if (quot && nq) HIP_OK(hipMemcpyAsync(quot, m.quot.p, 2 * nq, hipMemcpyDeviceToHost, S()));
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}
int dgx_da_decode_maps(dgx_da d, const uint8_t *rem, const uint8_t *qmask, const uint8_t *order,
                       const uint16_t *quot, size_t n_quot, uint32_t *nlist) {
    API_BEGIN
    ensure_ctx();
    if (!d || !rem || !qmask || !order || !nlist || (n_quot && !quot)) {
        dgx_set_error("null argument");
        return -1;
    }
    const DaDevice &a = d->get();
    const size_t ne = a.n_elem;
    LutMaps m;
    m.rem.alloc(8 * ne, S());
    m.qmask.alloc(ne, S());
    m.order.alloc(ne, S());
    m.quot.alloc(std::max<size_t>(n_quot, 1), S());
    m.nq = n_quot;
    DevBuf<uint32_t> out(8 * ne, S());
    if (ne) {
        HIP_OK(hipMemcpyAsync(m.rem.p, rem, 8 * ne, hipMemcpyHostToDevice, S()));
        HIP_OK(hipMemcpyAsync(m.qmask.p, qmask, ne, hipMemcpyHostToDevice, S()));
        HIP_OK(hipMemcpyAsync(m.order.p, order, ne, hipMemcpyHostToDevice, S()));
    }
    if (n_quot) HIP_OK(hipMemcpyAsync(m.quot.p, quot, 2 * n_quot, hipMemcpyHostToDevice, S()));
    da_lut_decode(a, m, out.p, S());
    if (ne) HIP_OK(hipMemcpyAsync(nlist, out.p, 32 * ne, hipMemcpyDeviceToHost, S()));
    HIP_OK(hipStreamSynchronize(S()));
    API_END
}

int dgx_da_matvec(dgx_da d, const double *x, double *y, double lap_fac, double mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!d || !x || !y) { dgx_set_error("matvec: null argument"); return -1; }
    da_matvec_f64(d->get(), x, y, lap_fac, mass_fac, S());
    API_END
}
int dgx_da_matvec_noalloc(dgx_da d, const double *x, double *y, double lap_fac, double mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!d || !x || !y) { dgx_set_error("matvec: null argument"); return -1; }
    da_matvec_f64_noalloc(d->get(), x, y, lap_fac, mass_fac, S());
    API_END
}
int dgx_da_matvec_f32(dgx_da d, const float *x, float *y, float lap_fac, float mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!d || !x || !y) { dgx_set_error("matvec: null argument"); return -1; }
    da_matvec_f32(d->get(), x, y, lap_fac, mass_fac, S());
    API_END
}
int dgx_da_diagonal(dgx_da d, double *diag, double lap_fac, double mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!d || !diag) { dgx_set_error("diagonal: null argument"); return -1; }
    da_diagonal_f64(d->get(), diag, lap_fac, mass_fac, S());
    API_END
}
int dgx_da_matvec_dirichlet(dgx_da d, const double *x, double *y, double lap_fac, double mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!d || !x || !y) { dgx_set_error("matvec: null argument"); return -1; }
    da_matvec_dirichlet_f64(d->get(), x, y, lap_fac, mass_fac, S());
    API_END
}

int dgx_mg_create(dgx_octree finest, int nlevels, double lap_fac, double mass_fac, dgx_mg *out) {
    API_BEGIN
    ensure_ctx();
    if (!finest || !out) { dgx_set_error("DAMG: null argument"); return -1; }
    if (nlevels < 1 || nlevels > 30) { dgx_set_error("DAMG: nlevels %d out of range", nlevels); return -1; }
    std::unique_ptr<dgx_mg_s> m(new dgx_mg_s());
    m->mg = mg_build(finest->keys.p, finest->n, finest->md, finest->curve, nlevels, lap_fac, mass_fac, S());
    m->finest.borrowed = m->mg->lev[0].da.get();
    *out = m.release();
    API_END
}
int dgx_mg_set_operator(dgx_mg m, double lap_fac, double mass_fac) {
    API_BEGIN
    ensure_ctx();
    if (!m) { dgx_set_error("DAMG: null argument"); return -1; }
    mg_set_operator(*m->mg, lap_fac, mass_fac, S());
    API_END
}
int dgx_mg_info(dgx_mg m, int *nlevels, uint64_t *octants, uint64_t *nodes) {
    if (!m) { dgx_set_error("null argument"); return -1; }
    const auto &lev = m->mg->lev;
    if (nlevels) *nlevels = (int)lev.size();
    for (size_t i = 0; i < lev.size() && i < 32; ++i) {
        if (octants) octants[i] = lev[i].n_oct;
        if (nodes) nodes[i] = lev[i].da->n_nodes;
    }
    return 0;
}
int dgx_mg_coarse_info(dgx_mg m, uint64_t out[6]) {
    if (!m || !out) { dgx_set_error("null argument"); return -1; }
    const CoarseSolver &c = *m->mg->coarse;
    out[0] = c.ni;
    out[1] = c.nfronts;
    out[2] = c.levels;
    out[3] = c.plan.max_m;
    out[4] = c.plan.max_s;
    out[5] = c.plan.front_entries;
    return 0;
}
int dgx_mg_level_da(dgx_mg m, int level, dgx_da *da) {
    if (!m || !da) { dgx_set_error("null argument"); return -1; }
    const int nl = (int)m->mg->lev.size();
    if (level < 0 || level >= nl) { dgx_set_error("DAMG: level %d out of range (%d levels)", level, nl); return -1; }
    if (m->levels.empty())
        for (int i = 0; i < nl; ++i) {
            m->levels.emplace_back(new dgx_da_s());
            m->levels.back()->borrowed = m->mg->lev[i].da.get();
        }
    *da = m->levels[level].get();
    return 0;
}
int dgx_mg_finest_da(dgx_mg m, dgx_da *da) {
    if (!m || !da) { dgx_set_error("null argument"); return -1; }
    *da = &m->finest;
    return 0;
}
int dgx_mg_rhs(dgx_mg m, const double *f, double *b) {
    API_BEGIN
    ensure_ctx();
    if (!m || !f || !b) { dgx_set_error("rhs: null argument"); return -1; }
    mg_rhs_from_nodal(*m->mg, f, b, S());
    API_END
}
int dgx_mg_restrict(dgx_mg m, int level, const double *fine, double *coarse) {
    API_BEGIN
    ensure_ctx();
    if (!m || !fine || !coarse) { dgx_set_error("restrict: null argument"); return -1; }
    const int nl = (int)m->mg->lev.size();
    if (level < 0 || level + 1 >= nl) { dgx_set_error("restrict: level %d has no coarser level (%d levels)", level, nl); return -1; }
    mg_restrict(*m->mg, (size_t)level, fine, coarse, S());
    API_END
}
int dgx_mg_prolong(dgx_mg m, int level, const double *coarse, double *fine) {
    API_BEGIN
    ensure_ctx();
    if (!m || !fine || !coarse) { dgx_set_error("prolong: null argument"); return -1; }
    const int nl = (int)m->mg->lev.size();
    if (level < 0 || level + 1 >= nl) { dgx_set_error("prolong: level %d has no coarser level (%d levels)", level, nl); return -1; }
    mg_prolong(*m->mg, (size_t)level, coarse, fine, S());
    API_END
}
int dgx_mg_solve(dgx_mg m, const double *b, double *u, const dgx_solve_opts *opts, dgx_solve_stats *stats) {
    API_BEGIN
    ensure_ctx();
    if (!m || !b || !u) { dgx_set_error("solve: null argument"); return -1; }
    SolveOpts o;
    if (opts) {
        o.fp32_vcycle = opts->fp32_vcycle;
        o.pre = opts->pre;
        o.post = opts->post;
        o.omega = opts->omega;
        o.rtol = opts->rtol;
        o.maxit = opts->maxit;
        o.use_graph = opts->use_graph;
    }
    if (o.pre < 1 || o.post < 0 || o.maxit < 1) { dgx_set_error("solve: bad options"); return -1; }
    SolveStats st;
    mg_solve(*m->mg, b, u, o, &st, S());
    if (stats) {
        stats->iterations = st.iterations;
        stats->rel_residual = st.rel_residual;
        stats->seconds = st.seconds;
    }
    API_END
}
int dgx_mg_free(dgx_mg m) {
    API_BEGIN
    if (m) {
        if (g_ctx.ready) HIP_OK(hipStreamSynchronize(S()));
        delete m;
    }
    API_END
}

int dgx_da_free(dgx_da d) {
    API_BEGIN
    if (d && d->borrowed) {
        dgx_set_error("this DA belongs to a multigrid hierarchy (free it with dgx_mg_free)");
        return -1;
    }
    if (d) {
        if (g_ctx.ready) HIP_OK(hipStreamSynchronize(S()));
        delete d;
    }
    API_END
}

}  // extern "C"
