// dgx_mg.hip -- geometric multigrid on the octree hierarchy and the FP64 CG
// driver (one rank, gfx950).
//
// Reference structure: DAMGCreateAndSetDA (src/omg/blockDiag.cpp:1312-2190)
// coarsens the finest octree level by level (coarsenOctree, coarsen.cpp:53-104)
// and balances each coarse octree, stopping early below 500 octants
// (:1384-1425); each level gets its own DA.  DAMGSetKSP/DAMGSolve (:780-1170)
// run PETSc CG preconditioned by a multiplicative V-cycle with Richardson +
// Jacobi smoothing and an LU coarse solve (bin/optionsFBM:26-57).  PETSc is not
// part of this build; the same method is implemented here directly:
//   * operator: homogeneous Dirichlet -lap+I (DirichletJacobianMatMult,
//     omgJac.C:958-1057): boundary rows identity, boundary columns dropped;
//   * transfer: P interpolates the coarse finite-element function at the fine
//     nodes (trilinear, with the coarse element's hanging-node constraints),
//     R = P^T; Dirichlet nodes are excluded from both;
//   * smoother: damped Jacobi x += omega D^-1 (b - A x), nu1 pre / nu2 post;
//   * coarsest level: dense inverse of the Dirichlet operator in FP64;
//   * outer: flexible preconditioned CG in FP64 (Polak-Ribiere beta), so the
//     V-cycle may run in FP32 while residuals, dot products, updates and the
//     coarse solve stay FP64.
#include "dgx_mg.h"


#include "../../include/dendro_mi355x.h"  // dgx_last_error

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <string>
#include <thread>

#include <chrono>
#include <mutex>
#include <cmath>

namespace {

constexpr int B = 256;
constexpr uint32_t NONE = 0xffffffffu;

// ------------------------------------------------------------- coarsen ----
__global__ void k_coarsen_flags(const Key *k, size_t n, uint32_t *keep, uint32_t *start) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t L = key_level(k[i]);
    uint32_t st = 0;
    if (L >= 1 && key_digit(k[i], L) == 0 && i + 7 < n && key_level(k[i + 7]) == L &&
        key_eq(key_parent(k[i + 7]), key_parent(k[i])))
        st = 1;
    start[i] = st;
    keep[i] = 1;
}
__global__ void k_coarsen_mark(const uint32_t *start, size_t n, uint32_t *keep) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !start[i]) return;
    for (int j = 1; j < 8; ++j) keep[i + j] = 0;
}
__global__ void k_coarsen_emit(const Key *k, const uint32_t *keep, const uint32_t *start, const uint32_t *pos, size_t n,
                               Key *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    out[pos[i]] = start[i] ? key_parent(k[i]) : k[i];
}

// ------------------------------------------------------------ transfer ----
__device__ inline bool mlb(const Key *a, size_t n, const Key &k, size_t &idx) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key_le(a[mid], k)) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) return false;
    idx = lo - 1;
    return true;
}

// One row of P per fine node: interpolate the coarse FE function at the node.
__global__ void __launch_bounds__(B) k_prolong_rows(uint32_t nf, const uint32_t *n2l_f, const uint32_t *xf,
                                                    const uint32_t *yf, const uint32_t *zf, const uint8_t *bdy_f,
                                                    const Key *ckey, size_t nce, const uint32_t *xc,
                                                    const uint32_t *yc, const uint32_t *zc, const uint8_t *cflags,
                                                    const uint32_t *cconn, const uint16_t *cmeta,
                                                    const uint8_t *bdy_c, uint32_t md, SfcPacked tab,
                                                    uint32_t *pcol, double *pval, uint32_t *err) {
    __shared__ uint8_t s_rc[192];
    for (int i = threadIdx.x; i < 192; i += blockDim.x) s_rc[i] = tab.rc[i];
    __syncthreads();
    const uint32_t n = blockIdx.x * B + threadIdx.x;
    if (n >= nf) return;
    for (int k = 0; k < 8; ++k) {
        pcol[(size_t)k * nf + n] = NONE;
        pval[(size_t)k * nf + n] = 0.0;
    }
    if (bdy_f[n]) return;
    const uint32_t l = n2l_f[n], D = 1u << (md - 1);
    const uint32_t X = xf[l], Y = yf[l], Z = zf[l];
    const uint32_t qx = X == D ? D - 1 : X, qy = Y == D ? D - 1 : Y, qz = Z == D ? D - 1 : Z;
    size_t e;
    if (!mlb(ckey, nce, dev_encode(s_rc, qx, qy, qz, md, md), e)) {
        atomicOr(err, 1u);
        return;
    }
    const uint32_t lc = cflags[e] & 31u, sc = 1u << (md - lc);
    const uint32_t ax = xc[e], ay = yc[e], az = zc[e];
    if (X < ax || Y < ay || Z < az || X > ax + sc || Y > ay + sc || Z > az + sc) {
        atomicOr(err, 2u);
        return;
    }
    // local coordinates in halves: 0, 1 (midpoint) or 2
    const uint64_t dx2 = 2ull * (X - ax), dy2 = 2ull * (Y - ay), dz2 = 2ull * (Z - az);
    if ((dx2 % sc) || (dy2 % sc) || (dz2 % sc)) {
        atomicOr(err, 4u);
        return;
    }
    const uint32_t tx = (uint32_t)(dx2 / sc), ty = (uint32_t)(dy2 / sc), tz = (uint32_t)(dz2 / sc);
    const uint32_t mt = cmeta[e], mask = mt & 255u, c = (mt >> 8) & 7u;
    double w[8];
    for (int k = 0; k < 8; ++k) w[k] = 0.0;
    for (uint32_t a = 0; a < 8; ++a) {  // trilinear weight of coarse vertex a
        const uint32_t bx = a & 1, by = (a >> 1) & 1, bz = (a >> 2) & 1;
        const double wx = bx ? tx * 0.5 : 1.0 - tx * 0.5, wy = by ? ty * 0.5 : 1.0 - ty * 0.5,
                     wz = bz ? tz * 0.5 : 1.0 - tz * 0.5;
        const double wa = wx * wy * wz;
        if (wa == 0.0) continue;
        if (!((mask >> a) & 1u)) {
            w[a] += wa;
            continue;
        }
        const uint32_t F = a ^ c;
        const double share = wa / (double)(1u << __popc(F));
        for (uint32_t f = 0; f < 8; ++f)
            if ((f & ~F) == 0) w[a ^ f] += share;
    }
    int o = 0;
    for (int k = 0; k < 8; ++k) {
        if (w[k] == 0.0) continue;
        const uint32_t col = cconn[(size_t)k * nce + e];
        if (bdy_c[col]) continue;
        pcol[(size_t)o * nf + n] = col;
        pval[(size_t)o * nf + n] = w[k];
        ++o;
    }
}
// (coarse node, fine node) pairs of P's entries, row n's at P's own row
// offsets poff[n] (the rows are filled from slot 0 and NONE ends them)
__global__ void k_prolong_keys(const uint32_t *pcol, const uint32_t *poff, uint32_t nf, Key *keys, uint32_t *vals) {
    const uint32_t n = blockIdx.x * B + threadIdx.x;
    if (n >= nf) return;
    const uint32_t base = poff[n], nloc = poff[n + 1] - base;
    for (uint32_t k = 0; k < nloc; ++k) {
        keys[base + k] = Key{(uint64_t)pcol[(size_t)k * nf + n], (uint64_t)n};
        vals[base + k] = k * nf + n;
    }
}
__global__ void k_restrict_csr(const Key *keys, const uint32_t *vals, uint32_t nnz, const double *pval, uint32_t *rcol,
                               double *rval, float *rval32, uint32_t *rowcnt) {
    uint32_t p = blockIdx.x * B + threadIdx.x;
    if (p >= nnz) return;
    rcol[p] = (uint32_t)keys[p].lo;
    rval[p] = pval[vals[p]];
    rval32[p] = (float)pval[vals[p]];
    atomicAdd(&rowcnt[(uint32_t)keys[p].hi], 1u);
}
__global__ void k_to_f32(const double *a, size_t n, float *b) {
    size_t i = (size_t)blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_dinv(const double *diag, const uint8_t *bdy, uint32_t n, double *dinv, float *dinv32) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    double v = bdy[i] ? 1.0 : 1.0 / diag[i];
    dinv[i] = v;
    dinv32[i] = (float)v;
}

// Gauss-Jordan inversion of an SPD matrix, one pivot per launch (no pivoting).
// Column-major lower triangle (row >= col) -> full symmetric matrix.

// -------------------------------------------------------- vector kernels ----
template <typename T>
__global__ void k_jacobi_first(const T *b, const T *dinv, T omega, uint32_t n, T *x) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) x[i] = omega * dinv[i] * b[i];
}
// bc = R rf: 8 lanes per coarse row, strided over the row's entries, then a
// butterfly sum inside the 8-lane group.
template <typename T>
__global__ void __launch_bounds__(B) k_restrict(const uint32_t *__restrict__ roff, const uint32_t *__restrict__ rcol,
                                                const T *__restrict__ rval, uint32_t nc, const T *__restrict__ rf,
                                                T *__restrict__ bc) {
    const uint32_t t = blockIdx.x * B + threadIdx.x, i = t >> 3, sub = t & 7u;
    T acc = T(0);
    if (i < nc) {
        // four entries per lane per round: their columns and weights, then
        // their gathers, are in flight together; the summation order is the
        // entry order
        const uint32_t p1 = roff[i + 1];
        for (uint32_t pb = roff[i] + sub; pb < p1; pb += 32) {
            uint32_t c[4];
            T v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t p = pb + 8 * k;
                c[k] = p < p1 ? rcol[p] : 0u;
                v[k] = p < p1 ? rval[p] : T(0);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (pb + 8 * k < p1) acc += v[k] * rf[c[k]];
        }
    }
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 1, 64);
    if (i < nc && sub == 0) bc[i] = acc;
}
// xf += P xc, P in CSR over fine nodes
template <typename T>
__global__ void __launch_bounds__(B) k_prolong_add(const uint32_t *__restrict__ poff, const uint32_t *__restrict__ pcol,
                                                   const T *__restrict__ pval, uint32_t nf, const T *__restrict__ xc,
                                                   T *__restrict__ xf) {
    const uint32_t n = blockIdx.x * B + threadIdx.x;
    if (n >= nf) return;
    // a row has at most 8 entries: load them all, then gather all
    const uint32_t p0 = poff[n], p1 = poff[n + 1];
    uint32_t c[8];
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        c[k] = p0 + k < p1 ? pcol[p0 + k] : 0u;
        v[k] = p0 + k < p1 ? pval[p0 + k] : T(0);
    }
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (p0 + k < p1) acc += v[k] * xc[c[k]];
    xf[n] += acc;
}
__global__ void k_ell_count(const uint32_t *ecol, uint32_t nf, uint32_t *cnt) {
    const uint32_t n = blockIdx.x * B + threadIdx.x;
    if (n >= nf) return;
    uint32_t c = 0;
    while (c < 8 && ecol[(size_t)c * nf + n] != NONE) ++c;
    cnt[n] = c;
}
__global__ void k_ell_to_csr(const uint32_t *ecol, const double *eval, uint32_t nf, const uint32_t *off,
                             uint32_t *pcol, double *pval, float *pval32) {
    const uint32_t n = blockIdx.x * B + threadIdx.x;
    if (n >= nf) return;
    uint32_t o = off[n];
    for (int k = 0; k < 8; ++k) {
        const uint32_t c = ecol[(size_t)k * nf + n];
        if (c == NONE) break;
        pcol[o] = c;
        pval[o] = eval[(size_t)k * nf + n];
        pval32[o] = (float)eval[(size_t)k * nf + n];
        ++o;
    }
}
__global__ void k_f64_to_f32(const double *a, uint32_t n, float *b) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_f32_to_f64(const float *a, uint32_t n, double *b) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
// Two-stage deterministic dot products: per-block partial sums, then one block.
// Per block: the sum over its grid-stride share, DOT_U pairs per thread
// loaded ahead of their products (a fixed order, so the result is
// deterministic).  A one-kernel form, in which the block that finishes last
// sums the partials, was slower: each block's release and acquire at device
// scope write back and invalidate its L2.
constexpr int DOT_U = 4;
__global__ void k_dot_partial(const double *a, const double *b, uint32_t n, double *part) {
    __shared__ double sh[B / 64];
    double acc = 0.0;
    for (uint32_t i0 = blockIdx.x * B * DOT_U + threadIdx.x; i0 < n; i0 += gridDim.x * B * DOT_U) {
        double va[DOT_U], vb[DOT_U];
#pragma unroll
        for (int u = 0; u < DOT_U; ++u) {
            const uint32_t i = i0 + u * B;
            va[u] = i < n ? a[i] : 0.0;
            vb[u] = i < n ? b[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < DOT_U; ++u) acc += va[u] * vb[u];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < B / 64; ++w) t += sh[w];
        part[blockIdx.x] = t;
    }
}
__global__ void k_dot_final(const double *part, uint32_t np, double *out) {
    __shared__ double sh[B / 64];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += B) acc += part[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < B / 64; ++w) t += sh[w];
        *out = t;
    }
}
// CG updates with scalars held in device memory (s[0]=alpha numerator etc.)
__global__ void k_cg_xr(const double *p, const double *q, const double *sc, uint32_t n, double *x, double *r,
                        double *rold) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double alpha = sc[0] / sc[1];
    rold[i] = r[i];
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
}
__global__ void k_cg_p(const double *z, const double *sc, uint32_t n, double *p) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i >= n) return;
    const double beta = (sc[2] - sc[3]) / sc[0];  // (z.r - z.r_old) / (z_old . r_old)
    p[i] = z[i] + beta * p[i];
}
__global__ void k_mask_bdy(const uint8_t *bdy, uint32_t n, double *v) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n && bdy[i]) v[i] = 0.0;
}

constexpr uint32_t DOT_BLOCKS = 1024;

}  // namespace

// ---------------------------------------------------------------- build ----
DevBuf<Key> coarsen_build(const Key *in, size_t n, hipStream_t s) {
    DevBuf<Key> out;
    if (n < 8) {
        out.alloc(n, s);
        if (n) HIP_OK(hipMemcpyAsync(out.p, in, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
        return out;
    }
    DevBuf<uint32_t> keep(n, s), start(n, s), pos(n, s);
    k_coarsen_flags<<<grid_for(n, B), B, 0, s>>>(in, n, keep.p, start.p);
    k_coarsen_mark<<<grid_for(n, B), B, 0, s>>>(start.p, n, keep.p);
    uint64_t m = scan_exclusive_u32(keep.p, pos.p, n, s);
    out.alloc(m, s);
    k_coarsen_emit<<<grid_for(n, B), B, 0, s>>>(in, keep.p, start.p, pos.p, n, out.p);
    HIP_OK(hipGetLastError());
    return out;
}

static void build_transfer(MgLevel &fine, MgLevel &coarse, uint32_t md0, int curve, hipStream_t s) {
    const DaDevice &f = *fine.da, &c = *coarse.da;
    Transfer &t = coarse.to_finer;
    const uint32_t nf = f.n_nodes;
    t.nf = nf;
    t.nc = c.n_nodes;
    DevBuf<uint32_t> ecol(8 * (size_t)nf, s);  // ELL [8][nf], NONE-terminated rows
    DevBuf<double> eval(8 * (size_t)nf, s);
    DevBuf<uint32_t> err(1, s);
    HIP_OK(hipMemsetAsync(err.p, 0, 4, s));
    k_prolong_rows<<<grid_for(nf, B), B, 0, s>>>(nf, f.n2l.p, f.x.p, f.y.p, f.z.p, f.bdy_node.p, c.key.p, c.n_elem,
                                                c.x.p, c.y.p, c.z.p, c.flags.p, c.conn.p, c.meta.p, c.bdy_node.p,
                                                md0 + 1, make_packed(curve), ecol.p, eval.p, err.p);
    HIP_OK(hipGetLastError());
    uint64_t pn = 0;
    {
        DevBuf<uint32_t> cnt(nf, s);
        k_ell_count<<<grid_for(nf, B), B, 0, s>>>(ecol.p, nf, cnt.p);
        t.poff.alloc((size_t)nf + 1, s);
        pn = scan_exclusive_u32(cnt.p, t.poff.p, nf, s);
        set_u32_async(t.poff.p + nf, (uint32_t)pn, s);
        t.pcol.alloc(pn, s);
        t.pval.alloc(pn, s);
        t.pval32.alloc(pn, s);
        k_ell_to_csr<<<grid_for(nf, B), B, 0, s>>>(ecol.p, eval.p, nf, t.poff.p, t.pcol.p, t.pval.p, t.pval32.p);
        HIP_OK(hipGetLastError());
    }
    // transpose into CSR over coarse nodes
    DevBuf<Key> keys(pn, s);
    DevBuf<uint32_t> vals(pn, s);
    k_prolong_keys<<<grid_for(nf, B), B, 0, s>>>(ecol.p, t.poff.p, nf, keys.p, vals.p);
    const uint32_t hcnt = (uint32_t)pn;
    uint32_t herr = 0;
    dgx_read_back(&herr, err.p, 4, s);
    if (herr) {
        dgx_set_error("multigrid: interpolation lookup failed (code %u)", herr);
        throw DgxError();
    }
    t.nnz = hcnt;
    radix_sort_pairs(keys.p, vals.p, hcnt, s, nf, t.nc);  // (coarse node, fine node)
    t.rcol.alloc(hcnt, s);
    t.rval.alloc(hcnt, s);
    t.rval32.alloc(hcnt, s);
    DevBuf<uint32_t> rowcnt(t.nc, s);
    HIP_OK(hipMemsetAsync(rowcnt.p, 0, t.nc * 4, s));
    if (hcnt)
        k_restrict_csr<<<grid_for(hcnt, B), B, 0, s>>>(keys.p, vals.p, hcnt, eval.p, t.rcol.p, t.rval.p,
                                                      t.rval32.p, rowcnt.p);
    t.roff.alloc(t.nc + 1, s);
    uint64_t tot = scan_exclusive_u32(rowcnt.p, t.roff.p, t.nc, s);
    set_u32_async(t.roff.p + t.nc, (uint32_t)tot, s);
    HIP_OK(hipStreamSynchronize(s));
}

static void build_level_diag(MgLevel &L, double lap, double mass, hipStream_t s) {
    const DaDevice &d = *L.da;
    DevBuf<double> diag(d.n_nodes, s);
    da_diagonal_f64(d, diag.p, lap, mass, s);
    L.dinv.alloc(d.n_nodes, s);
    L.dinv32.alloc(d.n_nodes, s);
    k_dinv<<<grid_for(d.n_nodes, B), B, 0, s>>>(diag.p, d.bdy_node.p, d.n_nodes, L.dinv.p, L.dinv32.p);
    HIP_OK(hipGetLastError());
}

// Coarsest level: symbolic analysis (nested dissection, host) and numeric
// factorisation (device) of the sparse FP64 direct solver (dgx_direct.hip).
static double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}
static void build_coarse(MgHierarchy &mg, hipStream_t s, uint32_t dense, double *ms_analyse = nullptr,
                         double *ms_factor = nullptr) {
    const DaDevice &d = *mg.lev.back().da;
    auto t0 = std::chrono::steady_clock::now();
    mg.coarse = coarse_analyse(d, s, dense);
    if (ms_analyse) {
        HIP_OK(hipStreamSynchronize(s));
        *ms_analyse = ms_since(t0);
    }
    t0 = std::chrono::steady_clock::now();
    coarse_factor(*mg.coarse, d, mg.lap_fac, mg.mass_fac, s);
    if (ms_factor) {
        HIP_OK(hipStreamSynchronize(s));
        *ms_factor = ms_since(t0);
    }
}

namespace {

// The stream on which a calling thread's set-up helper works: created once per
// thread at default priority and kept for the process, because device buffers
// built on it are freed on it when the hierarchy goes.
hipStream_t helper_stream() {
    thread_local hipStream_t h = nullptr;
    if (!h) HIP_OK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    return h;
}

// Runs set-up tasks in order on a second host thread and stream, so that the
// DA of one level is built while the next level is coarsened and balanced.
// Device memory is charged to the caller's allocation tag.  A failed task
// keeps its message; the tasks after it are skipped, and wait() rethrows it on
// the calling thread.  The destructor joins the thread on every exit path.
// Without a thread (inline), submit runs the task at once on the caller's
// stream and synchronises it.
class SetupHelper {
  public:
    SetupHelper(hipStream_t hs, hipStream_t caller, bool threaded)
        : hs_(threaded ? hs : caller), tag_(dgx_alloc_get_tag()), dev_(current_device()), inline_(!threaded) {
        if (threaded) th_ = std::thread([this] { run(); });
    }
    ~SetupHelper() {
        if (inline_) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    // returns the task's index
    size_t submit(std::function<void(hipStream_t)> f) {
        if (inline_) {
            f(hs_);
            HIP_OK(hipStreamSynchronize(hs_));
            ++done_;
            return queued_++;
        }
        std::lock_guard<std::mutex> lk(mu_);
        todo_.push_back(std::move(f));
        const size_t k = queued_++;
        cv_.notify_all();
        return k;
    }
    // waits until tasks 0..k have run; rethrows a failure
    void wait(size_t k) {
        if (inline_) return;  // ran in submit; a failure was thrown there
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return done_ > k; });
        if (failed_) {
            dgx_set_error("%s", err_.c_str());
            throw DgxError();
        }
    }
    void wait_all() {
        size_t q;
        {
            std::lock_guard<std::mutex> lk(mu_);
            q = queued_;
        }
        if (q) wait(q - 1);
    }

  private:
    static int current_device() {
        int d = 0;
        HIP_OK(hipGetDevice(&d));
        return d;
    }
    void run() {
        (void)hipSetDevice(dev_);  // the calling thread's device (the current device is per thread)
        dgx_alloc_set_tag(tag_);
        for (;;) {
            std::function<void(hipStream_t)> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !todo_.empty(); });
                if (todo_.empty()) break;
                f = std::move(todo_.front());
                todo_.pop_front();
            }
            bool skip;
            {
                std::lock_guard<std::mutex> lk(mu_);
                skip = failed_;
            }
            std::string msg;
            bool bad = false;
            if (!skip) {
                try {
                    f(hs_);
                    HIP_OK(hipStreamSynchronize(hs_));
                } catch (DgxError &) {
                    bad = true;
                    msg = dgx_last_error();
                } catch (std::exception &e) {
                    bad = true;
                    msg = e.what();
                }
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (bad && !failed_) {
                    failed_ = true;
                    err_ = msg;
                }
                ++done_;
            }
            cv_.notify_all();
        }
        (void)hipStreamSynchronize(hs_);
        dgx_alloc_set_tag(0);
    }
    hipStream_t hs_;
    uint64_t tag_;
    int dev_;
    bool inline_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void(hipStream_t)>> todo_;
    size_t queued_ = 0, done_ = 0;
    bool stop_ = false, failed_ = false;
    std::string err_;
    std::thread th_;  // started in the constructor body, once the members above exist
};

}  // namespace

std::unique_ptr<MgHierarchy> mg_build(const Key *finest, size_t n, uint32_t md, int curve, int nlevels,
                                      double lap_fac, double mass_fac, hipStream_t s, bool overlap, bool colour,
                                      uint32_t dense_nodes) {
    std::unique_ptr<MgHierarchy> mg(new MgHierarchy());
    mg->md = md;
    mg->curve = curve;
    mg->lap_fac = lap_fac;
    mg->mass_fac = mass_fac;
    if (nlevels < 1) nlevels = 1;
    // DGX_SETUP_TIMING=1: one line per set-up on stderr with the wall time of
    // each phase (diagnostics; the phases are timed on the thread that runs them)
    static const bool timing = [] {
        const char *e = getenv("DGX_SETUP_TIMING");
        return e && e[0] == '1';
    }();
    const auto t_start = std::chrono::steady_clock::now();
    double t_cb = 0, t_wait = 0, t_tr = 0, t_an = 0, t_fa = 0;
    std::atomic<double> t_da{0.0};
    mg->lev.reserve(32);  // the helper holds pointers to levels: no reallocation (at most 30 levels)
    mg->lev.emplace_back();
    MgLevel &L0 = mg->lev.back();
    L0.keys.alloc(n, s);
    HIP_OK(hipMemcpyAsync(L0.keys.p, finest, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipStreamSynchronize(s));  // the helper reads the keys on its own stream
    L0.n_oct = n;
    // Each level's DA and Jacobi diagonal are built by the helper while the
    // calling thread coarsens and balances the next level; the stopping rule
    // needs a level's node count only once the requested levels exist.
    SetupHelper helper(overlap ? helper_stream() : s, s, overlap);
    auto build_da = [&, md, curve, colour](MgLevel *L) {
        return helper.submit([L, md, curve, lap_fac, mass_fac, colour, &t_da](hipStream_t hs) {
            const auto t0 = std::chrono::steady_clock::now();
            L->da = da_build(L->keys.p, L->n_oct, md, curve, hs, colour);
            build_level_diag(*L, lap_fac, mass_fac, hs);
            if (timing) {
                HIP_OK(hipStreamSynchronize(hs));
                t_da = t_da + ms_since(t0);
            }
        });
    };
    size_t last = build_da(&L0);
    // nlevels is honoured (blockDiag.cpp:1384-1425).  Only a coarsest level
    // beyond the sparse direct solver's reach (kMaxDirectNodes) makes the
    // hierarchy coarsen further.
    for (int k = 1; k < 30; ++k) {
        if (k >= nlevels) {
            const auto tw = std::chrono::steady_clock::now();
            helper.wait(last);
            t_wait += ms_since(tw);
            if (mg->lev.back().da->n_nodes <= kMaxDirectNodes) break;
        }
        MgLevel &prev = mg->lev.back();
        const auto tc = std::chrono::steady_clock::now();
        DevBuf<Key> c = coarsen_build(prev.keys.p, prev.n_oct, s);
        DevBuf<Key> b = balance_build(c.p, c.n, md, curve, s, nullptr);  // synchronises s
        t_cb += ms_since(tc);
        if (b.n >= prev.n_oct || b.n <= 8) break;  // nothing left to coarsen
        mg->lev.emplace_back();
        MgLevel &L = mg->lev.back();
        L.n_oct = b.n;
        L.keys = std::move(b);
        last = build_da(&L);
        if (L.n_oct < 500) break;  // blockDiag.cpp:1417-1420
    }
    {
        const auto tw = std::chrono::steady_clock::now();
        helper.wait_all();
        t_wait += ms_since(tw);
    }
    if (overlap) {
        // the coarse solver on the helper, the transfers here
        helper.submit([&mg, &t_an, &t_fa, dense_nodes](hipStream_t hs) {
            build_coarse(*mg, hs, dense_nodes, timing ? &t_an : nullptr, timing ? &t_fa : nullptr);
        });
        const auto tt = std::chrono::steady_clock::now();
        for (size_t k = 1; k < mg->lev.size(); ++k) build_transfer(mg->lev[k - 1], mg->lev[k], md, curve, s);
        if (timing) HIP_OK(hipStreamSynchronize(s));
        t_tr = ms_since(tt);
        helper.wait_all();
    } else {
        // One stream: the coarsest level's host analysis runs on a second host
        // thread (it makes no HIP calls) while this thread builds the
        // transfers on s; then the plan is uploaded and factorised on s.
        const DaDevice &dc = *mg->lev.back().da;
        const CoarseInputs in = coarse_download(dc, s);
        MfPlan plan;
        std::string perr;
        bool pfail = false;
        std::thread an([&] {
            const auto ta = std::chrono::steady_clock::now();
            try {
                plan = coarse_plan(in, dense_nodes);
            } catch (DgxError &) {
                pfail = true;
                perr = dgx_last_error();
            } catch (std::exception &e) {
                pfail = true;
                perr = e.what();
            }
            t_an = ms_since(ta);
        });
        struct Join {
            std::thread &t;
            ~Join() {
                if (t.joinable()) t.join();
            }
        } join{an};
        const auto tt = std::chrono::steady_clock::now();
        for (size_t k = 1; k < mg->lev.size(); ++k) build_transfer(mg->lev[k - 1], mg->lev[k], md, curve, s);
        if (timing) HIP_OK(hipStreamSynchronize(s));
        t_tr = ms_since(tt);
        an.join();
        if (pfail) {
            dgx_set_error("%s", perr.c_str());
            throw DgxError();
        }
        mg->coarse = coarse_upload(std::move(plan), in.n, in.E, s);
        const auto tf = std::chrono::steady_clock::now();
        coarse_factor(*mg->coarse, dc, mg->lap_fac, mg->mass_fac, s);
        if (timing) {
            HIP_OK(hipStreamSynchronize(s));
            t_fa = ms_since(tf);
        }
    }
    HIP_OK(hipStreamSynchronize(s));
    if (timing)
        fprintf(stderr,
                "mg_build: %zu octants, %zu levels, coarsest %u nodes; total %.2f ms: coarsen+balance %.2f, "
                "DA+diagonal %.2f (helper), waits %.2f, transfers %.2f, coarse analysis %.2f, factorisation %.2f\n",
                n, mg->lev.size(), mg->lev.back().da->n_nodes, ms_since(t_start), t_cb, t_da.load(), t_wait, t_tr,
                t_an, t_fa);
    return mg;
}

void mg_reset_graphs(MgHierarchy &mg);
void mg_set_operator(MgHierarchy &mg, double lap_fac, double mass_fac, hipStream_t s) {
    mg.lap_fac = lap_fac;
    mg.mass_fac = mass_fac;
    for (auto &L : mg.lev) build_level_diag(L, lap_fac, mass_fac, s);
    coarse_factor(*mg.coarse, *mg.lev.back().da, lap_fac, mass_fac, s);
    mg_reset_graphs(mg);
}

// ---------------------------------------------------------------- solve ----
namespace {

template <typename T>
struct Work {
    std::vector<DevBuf<T>> x, b, y;
};

template <typename T>
const T *dinv_of(const MgLevel &L);
template <>
const double *dinv_of<double>(const MgLevel &L) { return L.dinv.p; }
template <>
const float *dinv_of<float>(const MgLevel &L) { return L.dinv32.p; }
template <typename T>
const T *pval_of(const Transfer &t);
template <>
const double *pval_of<double>(const Transfer &t) { return t.pval.p; }
template <>
const float *pval_of<float>(const Transfer &t) { return t.pval32.p; }
template <typename T>
const T *rval_of(const Transfer &t);
template <>
const double *rval_of<double>(const Transfer &t) { return t.rval.p; }
template <>
const float *rval_of<float>(const Transfer &t) { return t.rval32.p; }


template <typename T>
void vcycle(MgHierarchy &mg, Work<T> &w, size_t k, const SolveOpts &o, hipStream_t s) {
    MgLevel &L = mg.lev[k];
    const uint32_t n = L.da->n_nodes;
    const unsigned g = grid_for(n, B);
    T *x = w.x[k].p, *b = w.b[k].p, *y = w.y[k].p;
    if (k + 1 == mg.lev.size()) {  // coarsest: sparse FP64 direct solve
        coarse_solve<T>(*mg.coarse, L.da->bdy_node.p, b, x, s);
        return;
    }
    const T om = (T)o.omega, lap = (T)mg.lap_fac, mass = (T)mg.mass_fac;
    const T *dinv = dinv_of<T>(L);
    // Jacobi sweeps x += omega D^-1 (b - A x) and the residual b - A x, each
    // one matvec with the update fused into its stores; the first sweep
    // starts from x = 0 and needs no matvec
    k_jacobi_first<T><<<g, B, 0, s>>>(b, dinv, om, n, x);
    for (int it = 1; it < o.pre; ++it) da_dirichlet_jacobi<T>(*L.da, x, b, dinv, om, lap, mass, y, s);
    da_dirichlet_residual<T>(*L.da, x, b, y, lap, mass, s);
    const Transfer &t = mg.lev[k + 1].to_finer;
    k_restrict<T><<<grid_for((size_t)t.nc * 8, B), B, 0, s>>>(t.roff.p, t.rcol.p, rval_of<T>(t), (uint32_t)t.nc, y, w.b[k + 1].p);
    vcycle<T>(mg, w, k + 1, o, s);
    k_prolong_add<T><<<g, B, 0, s>>>(t.poff.p, t.pcol.p, pval_of<T>(t), n, w.x[k + 1].p, x);
    for (int it = 0; it < o.post; ++it) da_dirichlet_jacobi<T>(*L.da, x, b, dinv, om, lap, mass, y, s);
}

template <typename T>
void ensure_work(MgHierarchy &mg, Work<T> &w, hipStream_t s) {
    if (w.x.size() == mg.lev.size()) return;
    w.x.clear();
    w.b.clear();
    w.y.clear();
    for (auto &L : mg.lev) {
        w.x.emplace_back(L.da->n_nodes, s);
        w.b.emplace_back(L.da->n_nodes, s);
        w.y.emplace_back(L.da->n_nodes, s);
    }
}

__global__ void k_copy(const double *a, uint32_t n, double *b) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void k_zero(uint32_t n, double *b) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) b[i] = 0.0;
}
__global__ void k_scal_copy(double *sc, int from, int to) {
    if (threadIdx.x == 0 && blockIdx.x == 0) sc[to] = sc[from];
}

struct Graph {
    hipGraphExec_t exec = nullptr;
    SolveOpts key{};
    bool valid = false;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        valid = false;
    }
    ~Graph() { reset(); }
};

std::mutex g_capture_mu;

bool same_opts(const SolveOpts &a, const SolveOpts &b) {
    return a.fp32_vcycle == b.fp32_vcycle && a.pre == b.pre && a.post == b.post && a.omega == b.omega;
}

}  // namespace

// CG vectors, device scalars and captured graphs, owned by the hierarchy so
// that every captured pointer stays valid for the hierarchy's lifetime.
// sc: [0] z.r  [1] p.q  [2] z_new.r_new  [3] z_new.r_old  [4] r.r  [5] b.b
struct MgWork {
    Work<double> w64;
    Work<float> w32;
    DevBuf<double> x, bvec, r, z, p, q, rold, part, sc;
    Graph g_precond, g_step, g_dir;
    hipStream_t last = nullptr;  // stream of the last solve
    // No replay may still be running when the executable graphs are destroyed.
    // Only that stream is waited for: a device-wide wait would break graph
    // captures running on other streams (request server).
    ~MgWork() {
        // when frees are redirected the owner has already waited for this work
        if (last && !dgx_alloc_free_stream_get()) (void)hipStreamSynchronize(last);
    }
};
MgHierarchy::~MgHierarchy() { delete work; }
void mg_mark_idle(MgHierarchy &mg) {
    if (mg.work) mg.work->last = nullptr;
}
void mg_reset_graphs(MgHierarchy &mg) {
    if (!mg.work) return;
    if (mg.work->last) HIP_OK(hipStreamSynchronize(mg.work->last));
    mg.work->g_precond.reset();
    mg.work->g_step.reset();
    mg.work->g_dir.reset();
}

namespace {

MgWork &work_of(MgHierarchy &mg, hipStream_t s) {
    if (!mg.work) {
        mg.work = new MgWork();
        const uint32_t n = mg.lev[0].da->n_nodes;
        MgWork &w = *mg.work;
        w.x.alloc(n, s);
        w.bvec.alloc(n, s);
        w.r.alloc(n, s);
        w.z.alloc(n, s);
        w.p.alloc(n, s);
        w.q.alloc(n, s);
        w.rold.alloc(n, s);
        w.part.alloc(DOT_BLOCKS, s);
        w.sc.alloc(8, s);
        ensure_work(mg, w.w64, s);
        ensure_work(mg, w.w32, s);
    }
    return *mg.work;
}

void dot_dev(const double *a, const double *b, uint32_t n, double *part, double *out, hipStream_t s) {
    k_dot_partial<<<DOT_BLOCKS, B, 0, s>>>(a, b, n, part);
    k_dot_final<<<1, B, 0, s>>>(part, DOT_BLOCKS, out);
}

// z = M r on the work vectors (V-cycle in FP32 or FP64).
void precond_body(MgHierarchy &mg, MgWork &w, const SolveOpts &o, hipStream_t s) {
    const uint32_t n = mg.lev[0].da->n_nodes;
    if (o.fp32_vcycle) {
        k_f64_to_f32<<<grid_for(n, B), B, 0, s>>>(w.r.p, n, w.w32.b[0].p);
        vcycle<float>(mg, w.w32, 0, o, s);
        k_f32_to_f64<<<grid_for(n, B), B, 0, s>>>(w.w32.x[0].p, n, w.z.p);
    } else {
        k_copy<<<grid_for(n, B), B, 0, s>>>(w.r.p, n, w.w64.b[0].p);
        vcycle<double>(mg, w.w64, 0, o, s);
        k_copy<<<grid_for(n, B), B, 0, s>>>(w.w64.x[0].p, n, w.z.p);
    }
}
// q = A p; p.q; x += alpha p; r -= alpha q; r.r
void step_body(MgHierarchy &mg, MgWork &w, hipStream_t s) {
    const DaDevice &d = *mg.lev[0].da;
    const uint32_t n = d.n_nodes;
    da_matvec_dirichlet_f64(d, w.p.p, w.q.p, mg.lap_fac, mg.mass_fac, s);
    dot_dev(w.p.p, w.q.p, n, w.part.p, w.sc.p + 1, s);
    k_cg_xr<<<grid_for(n, B), B, 0, s>>>(w.p.p, w.q.p, w.sc.p, n, w.x.p, w.r.p, w.rold.p);
    dot_dev(w.r.p, w.r.p, n, w.part.p, w.sc.p + 4, s);
}
// z = M r; z.r; z.r_old; p = z + beta p; z.r -> sc[0]
void direction_body(MgHierarchy &mg, MgWork &w, const SolveOpts &o, hipStream_t s) {
    const uint32_t n = mg.lev[0].da->n_nodes;
    precond_body(mg, w, o, s);
    dot_dev(w.z.p, w.r.p, n, w.part.p, w.sc.p + 2, s);
    dot_dev(w.z.p, w.rold.p, n, w.part.p, w.sc.p + 3, s);
    k_cg_p<<<grid_for(n, B), B, 0, s>>>(w.z.p, w.sc.p, n, w.p.p);
    k_scal_copy<<<1, 64, 0, s>>>(w.sc.p, 2, 0);
}

template <typename F>
void run_graph(Graph &g, const SolveOpts &o, bool use_graph, hipStream_t s, F body) {
    if (!use_graph) {
        body();
        return;
    }
    if (!g.valid || !same_opts(g.key, o)) {
        g.reset();
        // One capture at a time in the process: request-server slots solve on
        // their own threads and streams, and the capture and instantiation of
        // a graph then never run concurrently with another thread's.
        std::lock_guard<std::mutex> one_capture(g_capture_mu);
        hipGraph_t graph = nullptr;
        // the allocator queues no frees on s while it is captured
        dgx_alloc_capture_begin(s);
        struct Unmark {
            hipStream_t s;
            ~Unmark() { dgx_alloc_capture_end(s); }
        } unmark{s};
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(s, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        HIP_OK(hipStreamEndCapture(s, &graph));
        HIP_OK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
        HIP_OK(hipGraphDestroy(graph));
        g.key = o;
        g.valid = true;
    }
    HIP_OK(hipGraphLaunch(g.exec, s));
}

double read_scalar(const double *p, hipStream_t s) {
    double h;
    dgx_read_back(&h, p, sizeof(double), s);
    return h;
}

}  // namespace

void mg_vcycle_apply(MgHierarchy &mg, const double *r, double *z, const SolveOpts &o, hipStream_t s) {
    MgWork &w = work_of(mg, s);
    if (w.last && w.last != s) HIP_OK(hipStreamSynchronize(w.last));
    w.last = s;
    const uint32_t n = mg.lev[0].da->n_nodes;
    k_copy<<<grid_for(n, B), B, 0, s>>>(r, n, w.r.p);
    run_graph(w.g_precond, o, o.use_graph, s, [&] { precond_body(mg, w, o, s); });
    k_copy<<<grid_for(n, B), B, 0, s>>>(w.z.p, n, z);
    HIP_OK(hipGetLastError());
}

void mg_rhs_from_nodal(MgHierarchy &mg, const double *f, double *b, hipStream_t s) {
    const DaDevice &d = *mg.lev[0].da;
    da_matvec_f64(d, f, b, 0.0, 1.0, s);  // M f
    k_mask_bdy<<<grid_for(d.n_nodes, B), B, 0, s>>>(d.bdy_node.p, d.n_nodes, b);
    HIP_OK(hipGetLastError());
}

// coarse = R fine and fine = P coarse between level k (finer) and k + 1, on
// compact node vectors in FP64: the transfers the V-cycle applies
// (restrictMatVecType1 / prolongMatVecType1 with the boundary nodes
// suppressed, src/omg/restrictMatVec.cpp:486-619, prolongMatVec.cpp:154-292).
void mg_restrict(MgHierarchy &mg, size_t k, const double *fine, double *coarse, hipStream_t s) {
    const Transfer &t = mg.lev[k + 1].to_finer;
    if (!t.nc) return;
    k_restrict<double><<<grid_for(t.nc * 8, B), B, 0, s>>>(t.roff.p, t.rcol.p, t.rval.p, (uint32_t)t.nc, fine, coarse);
    HIP_OK(hipGetLastError());
}
void mg_prolong(MgHierarchy &mg, size_t k, const double *coarse, double *fine, hipStream_t s) {
    const Transfer &t = mg.lev[k + 1].to_finer;
    const uint32_t n = mg.lev[k].da->n_nodes;
    if (!n) return;
    k_zero<<<grid_for(n, B), B, 0, s>>>(n, fine);
    k_prolong_add<double><<<grid_for(n, B), B, 0, s>>>(t.poff.p, t.pcol.p, t.pval.p, n, coarse, fine);
    HIP_OK(hipGetLastError());
}

void mg_solve(MgHierarchy &mg, const double *b, double *xout, const SolveOpts &o, SolveStats *st, hipStream_t s) {
    auto t0 = std::chrono::steady_clock::now();
    MgWork &w = work_of(mg, s);
    if (w.last && w.last != s) HIP_OK(hipStreamSynchronize(w.last));  // work vectors are shared
    w.last = s;
    const DaDevice &d = *mg.lev[0].da;
    const uint32_t n = d.n_nodes;
    const unsigned g = grid_for(n, B);
    k_copy<<<g, B, 0, s>>>(b, n, w.bvec.p);
    k_zero<<<g, B, 0, s>>>(n, w.x.p);
    k_copy<<<g, B, 0, s>>>(w.bvec.p, n, w.r.p);
    dot_dev(w.bvec.p, w.bvec.p, n, w.part.p, w.sc.p + 5, s);
    const double bnorm = std::sqrt(read_scalar(w.sc.p + 5, s));
    SolveStats res;
    if (bnorm == 0.0) {
        k_zero<<<g, B, 0, s>>>(n, xout);
        if (st) *st = res;
        return;
    }
    run_graph(w.g_precond, o, o.use_graph, s, [&] { precond_body(mg, w, o, s); });
    k_copy<<<g, B, 0, s>>>(w.z.p, n, w.p.p);
    dot_dev(w.r.p, w.z.p, n, w.part.p, w.sc.p + 0, s);
    int it = 0;
    double rel = 1.0;
    for (it = 1; it <= o.maxit; ++it) {
        run_graph(w.g_step, o, o.use_graph, s, [&] { step_body(mg, w, s); });
        rel = std::sqrt(read_scalar(w.sc.p + 4, s)) / bnorm;
        if (rel <= o.rtol) break;
        run_graph(w.g_dir, o, o.use_graph, s, [&] { direction_body(mg, w, o, s); });
    }
    // true residual in FP64
    da_dirichlet_residual<double>(d, w.x.p, w.bvec.p, w.r.p, mg.lap_fac, mg.mass_fac, s);
    dot_dev(w.r.p, w.r.p, n, w.part.p, w.sc.p + 6, s);
    res.rel_residual = std::sqrt(read_scalar(w.sc.p + 6, s)) / bnorm;
    k_copy<<<g, B, 0, s>>>(w.x.p, n, xout);
    res.iterations = it > o.maxit ? o.maxit : it;
    HIP_OK(hipStreamSynchronize(s));
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = res;
}
