// dgx_matvec.hip -- element matvec of -lap + I on the octree DA (gfx950).
//
// Reference: Jacobian1MatMult (examples/src/backend/odaJac.C:275-366) applies,
// per element, out[idx_k] += sum_j (h/2 L[c][t][k][j] + h^3/8 M[c][t][k][j]) in[idx_j]
// with the type-2 stencils of handleType2Stencils.C:678-1000.  Those tables
// are C^T K C where K is the Q1 stiffness/mass on [-1,1]^3 and C interpolates
// the element's hanging vertices from its parent's corners: a vertex hanging on
// a parent edge takes the mean of the two edge ends, one hanging on a parent
// face the mean of the four face corners (checked against the reference tables
// to 8e-16 by the test suite).  The kernels apply C, K and C^T directly from
// the hanging mask in the element's child frame, with K diagonalised by a 3-D
// Haar butterfly, so no stencil table is read.  Node sums are accumulated per
// element block in LDS as fixed-point integers, which makes every result
// independent of the order of the accumulations.
#include "dgx_da.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

namespace {

constexpr int B = 256;

// Unscaled 1-D operators along the axis with vertex stride S:
//   mhat = [[2,1],[1,2]] (3 x mass of [-1,1]),  khat = [[1,-1],[-1,1]] (2 x stiffness)
template <typename T, int S>
__device__ __forceinline__ void op_mhat(const T *u, T *o) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & S) continue;
        T a = u[i], b = u[i + S];
        o[i] = fma(a, T(2), b);
        o[i + S] = fma(b, T(2), a);
    }
}
template <typename T, int S>
__device__ __forceinline__ void op_khat(const T *u, T *o) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & S) continue;
        T d = u[i] - u[i + S];
        o[i] = d;
        o[i + S] = -d;
    }
}
// v = g1 (khat mhat mhat + mhat khat mhat + mhat mhat khat) u + g2 mhat mhat mhat u
// (x fastest), i.e. the Q1 operator with g1 = lapFac h/36, g2 = massFac h^3/216.
template <typename T>
__device__ __forceinline__ void apply_q1_hat(const T *u, T g1, T g2, T *v) {
    T a[8], b[8], c[8], d[8], e[8], w[8], kb[8];
    op_mhat<T, 4>(u, a);
    op_mhat<T, 2>(a, b);
    op_khat<T, 2>(a, c);
    op_khat<T, 4>(u, d);
    op_mhat<T, 2>(d, e);
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = fma(g1, c[i] + e[i], g2 * b[i]);
    op_mhat<T, 1>(w, v);
    op_khat<T, 1>(b, kb);
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = fma(g1, kb[i], v[i]);
}

// One element in its child frame (vertex 0 = the parent corner it shares, 7 =
// the parent centre).  Hanging vertices can only be the edge midpoints 1,2,4
// (mean of vertices 0 and e) and the face centres 3,5,6 (mean of 0, f and
// f's two edge vertices).  The Q1 operator is invariant under the axis
// reflections that map any child to child 0, so no table and no child number
// is needed: r = C^T (g1 K + g2 M) C x with C fixed by the 6-bit mask.
template <typename T>
__device__ __forceinline__ void element_child_frame(uint32_t cm, T lap_fac, T mass_fac, const T *x, T *r) {
    const uint32_t mask = cm & 255u, lev = cm >> 8;
    const T h = (T)ldexp(1.0, -(int)lev);
    const T g1 = lap_fac * h * T(1.0 / 36.0), g2 = mass_fac * h * h * h * T(1.0 / 216.0);
    if (mask == 0) {
        apply_q1_hat<T>(x, g1, g2, r);
        return;
    }
    const bool h1 = mask & 2u, h2 = mask & 4u, h4 = mask & 16u, h3 = mask & 8u, h5 = mask & 32u, h6 = mask & 64u;
    T u[8], v[8];
    u[0] = x[0];
    u[7] = x[7];
    u[1] = h1 ? T(0.5) * (x[0] + x[1]) : x[1];
    u[2] = h2 ? T(0.5) * (x[0] + x[2]) : x[2];
    u[4] = h4 ? T(0.5) * (x[0] + x[4]) : x[4];
    u[3] = h3 ? T(0.25) * ((x[0] + x[3]) + (x[1] + x[2])) : x[3];
    u[5] = h5 ? T(0.25) * ((x[0] + x[5]) + (x[1] + x[4])) : x[5];
    u[6] = h6 ? T(0.25) * ((x[0] + x[6]) + (x[2] + x[4])) : x[6];
    apply_q1_hat<T>(u, g1, g2, v);
    const T q3 = h3 ? T(0.25) * v[3] : T(0), q5 = h5 ? T(0.25) * v[5] : T(0), q6 = h6 ? T(0.25) * v[6] : T(0);
    const T e1 = h1 ? T(0.5) * v[1] : T(0), e2 = h2 ? T(0.5) * v[2] : T(0), e4 = h4 ? T(0.5) * v[4] : T(0);
    r[0] = v[0] + (e1 + e2 + e4) + (q3 + q5 + q6);
    r[1] = (h1 ? e1 : v[1]) + q3 + q5;
    r[2] = (h2 ? e2 : v[2]) + q3 + q6;
    r[4] = (h4 ? e4 : v[4]) + q5 + q6;
    r[3] = h3 ? q3 : v[3];
    r[5] = h5 ? q5 : v[5];
    r[6] = h6 ? q6 : v[6];
    r[7] = v[7];
}

// 3-D Haar butterfly on the 8 vertex values (x fastest): after it, entry i
// holds the sum/difference combination with a difference along axis k iff bit
// k of i is set.  The 1-D Q1 operators are diagonal in that basis,
//   mhat = [[2,1],[1,2]] = H diag(3,1) H / 2,  khat = [[1,-1],[-1,1]] = H diag(0,2) H / 2,
// with H = [[1,1],[1,-1]], so g1 (khat mhat mhat + mhat khat mhat + mhat mhat khat)
// + g2 mhat mhat mhat = H3 diag(lam) H3 / 8 where lam depends only on the number
// nd of difference axes: nd=0: 27 g2, 1: 18 g1 + 9 g2, 2: 12 g1 + 3 g2, 3: 6 g1 + g2.
template <typename T>
__device__ __forceinline__ void haar3(T *u) {
#pragma unroll
    for (int s = 1; s < 8; s <<= 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i & s) continue;
            const T a = u[i], b = u[i + s];
            u[i] = a + b;
            u[i + s] = a - b;
        }
    }
}

constexpr int PER = 2;  // elements per thread (element block = PER * B)
constexpr uint32_t ID_MASK = 0x3fffffffu, BDY_BIT = 0x40000000u, SHARED_BIT = 0x80000000u;
enum { EPI_Y = 0, EPI_JACOBI = 1, EPI_RESID = 2 };

// Haar eigenvalues / 8 per level (h = 2^-level), lam[level][nd].
template <typename T>
__device__ __forceinline__ void init_lambda(T (*lam)[4], double lap_fac, double mass_fac) {
    if (threadIdx.x < 32) {
        const double h = ldexp(1.0, -(int)threadIdx.x);
        const double g1 = lap_fac * h * (1.0 / 36.0), g2 = mass_fac * h * h * h * (1.0 / 216.0);
        lam[threadIdx.x][0] = (T)(27.0 * g2 * 0.125);
        lam[threadIdx.x][1] = (T)((18.0 * g1 + 9.0 * g2) * 0.125);
        lam[threadIdx.x][2] = (T)((12.0 * g1 + 3.0 * g2) * 0.125);
        lam[threadIdx.x][3] = (T)((6.0 * g1 + g2) * 0.125);
    }
}

// ---------------------------------------------------------------------------
// Fixed-point node sums.  Each block adds its element contributions into the
// LDS accumulators as 64-bit integers q = round(v 2^E), with one exponent E
// per block.  Integer addition is exact and associative, so a node's sum does
// not depend on the order in which the atomics land: the matvec, and
// everything built on it, is deterministic.  E comes from an a-priori bound.
// With u = C x (|u| <= X = max |x| over the block's nodes, C averages) and
// A_e = g1 K + g2 M, every contribution satisfies
//   |r_j| <= 3.25 X ||A_e||_inf <= 3.25 X (24 g1 + 27 g2),
// since K's rows sum to 0 around the diagonal 12 g1, M >= 0 has row sums
// 27 g2, and 3.25 = 1 + 3/2 + 3/4 bounds the hanging transpose C^T.  g1, g2
// are largest at the block's coarsest element.  With X < 2^xe and that factor
// < 2^se, E = 50 - xe - se keeps every |q| < 2^50 and a node's sum of at most
// 512 x 8 terms below 2^62.  The quantisation step is 2^-E, at most 2^-48 of
// the bound.  A non-finite x value makes every row its block touches NaN
// (with atomics in floating point, only its neighbours' rows).
// ---------------------------------------------------------------------------
constexpr int kExpZero = -2000, kExpBad = 4000;

// |v| < 2^mag_exp(v); kExpZero for 0, kExpBad for inf / NaN.
__device__ __forceinline__ int mag_exp(double v) {
    const double a = fabs(v);
    if (!(a <= 1.7976931348623157e308)) return kExpBad;
    if (a == 0.0) return kExpZero;
    return __builtin_amdgcn_frexp_exp(a);
}

// Maximum over the 64 lanes of a wave (every lane active), by DPP row shifts
// and row broadcasts; the result is wave-uniform.
__device__ __forceinline__ int wave_max_i32(int v) {
    constexpr int lo = -2147483647 - 1;
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

struct FixScale {
    double s;  // 2^E
    int e;     // E
    bool bad;  // non-finite input: rows become NaN
};

// se[lv]: the bound factor 3.25 (24 g1 + 27 g2) of an element of level lv is
// below 2^se[lv] (threads 0..31 fill the table; a barrier publishes it).
__device__ __forceinline__ void init_level_exps(int *se, double lap_fac, double mass_fac) {
    if (threadIdx.x < 32) {
        const double h = ldexp(1.0, -(int)threadIdx.x);
        se[threadIdx.x] =
            mag_exp(3.25 * (24.0 * fabs(lap_fac) * h * (1.0 / 36.0) + 27.0 * fabs(mass_fac) * h * h * h * (1.0 / 216.0)));
    }
}

// E for a block whose node values satisfy |x| < 2^xe and whose elements'
// bound factors are below 2^se (wave-uniform inputs, scalar results).
__device__ __forceinline__ FixScale fix_scale(int xe, int se) {
    FixScale f;
    int e = 50 - xe - se;
    // 2^E must be a normal double; a block past the lower end would need a
    // bound above 2^1070, beyond any finite x and operator factors in use
    f.bad = xe >= kExpBad || se >= kExpBad || e < -1022;
    e = f.bad ? 0 : min(e, 1023);
    e = __builtin_amdgcn_readfirstlane(e);
    f.e = e;
    f.s = __longlong_as_double((long long)(e + 1023) << 52);
    f.bad = __builtin_amdgcn_readfirstlane((int)f.bad) != 0;
    return f;
}

// round(v 2^E) as a two's-complement integer: for |v 2^E| < 2^51 the sum
// v 2^E + 1.5 2^52 is rounded to an integer once and holds it in its low
// mantissa bits.
__device__ __forceinline__ unsigned long long to_fixed(double v, double s) {
    const double t = fma(v, s, 6755399441055744.0);
    return (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;
}
__device__ __forceinline__ double from_fixed(unsigned long long q, const FixScale &f) {
    return f.bad ? __builtin_nan("") : ldexp((double)(long long)q, -f.e);
}

// Block maxima of xe and se through one LDS pair per wave; the caller's next
// __syncthreads publishes them and block_scale() reads them.
__device__ __forceinline__ void publish_scale_inputs(int xe, int se, int (*wred)[2]) {
    xe = wave_max_i32(xe);
    se = wave_max_i32(se);
    if ((threadIdx.x & 63u) == 0) {
        wred[threadIdx.x >> 6][0] = xe;
        wred[threadIdx.x >> 6][1] = se;
    }
}
__device__ __forceinline__ FixScale block_scale(const int (*wred)[2]) {
    int xe = wred[0][0], se = wred[0][1];
#pragma unroll
    for (int w = 1; w < B / 64; ++w) {
        xe = max(xe, wred[w][0]);
        se = max(se, wred[w][1]);
    }
    return fix_scale(xe, se);
}

// The element operator in its child frame, in place on the 8 vertex values:
// r = C^T (g1 K + g2 M) C u with C fixed by the hanging bits of cm.  With
// hanging == false the interpolation is skipped (the mask must then be 0).
template <typename T>
__device__ __forceinline__ void element_apply(T *u, uint32_t cm, bool hanging, const T (*lam)[4]) {
    const uint32_t mask = cm & 255u, lev = cm >> 8;
    bool h1 = false, h2 = false, h4 = false, h3 = false, h5 = false, h6 = false;
    if (hanging) {
        // hanging vertices in the child frame: edge midpoints 1,2,4 (mean of
        // vertex 0 and the edge end), face centres 3,5,6 (mean of 0, the face
        // vertex and its two edge vertices)
        h1 = mask & 2u;
        h2 = mask & 4u;
        h4 = mask & 16u;
        h3 = mask & 8u;
        h5 = mask & 32u;
        h6 = mask & 64u;
        const T x0 = u[0], x1 = u[1], x2 = u[2], x4 = u[4];
        if (h1) u[1] = T(0.5) * (x0 + x1);
        if (h2) u[2] = T(0.5) * (x0 + x2);
        if (h4) u[4] = T(0.5) * (x0 + x4);
        if (h3) u[3] = T(0.25) * ((x0 + u[3]) + (x1 + x2));
        if (h5) u[5] = T(0.25) * ((x0 + u[5]) + (x1 + x4));
        if (h6) u[6] = T(0.25) * ((x0 + u[6]) + (x2 + x4));
    }
    haar3<T>(u);
    const T l0 = lam[lev][0], l1 = lam[lev][1], l2 = lam[lev][2], l3 = lam[lev][3];
    u[0] *= l0;
    u[1] *= l1;
    u[2] *= l1;
    u[4] *= l1;
    u[3] *= l2;
    u[5] *= l2;
    u[6] *= l2;
    u[7] *= l3;
    haar3<T>(u);
    if (hanging) {
        const T q3 = h3 ? T(0.25) * u[3] : T(0), q5 = h5 ? T(0.25) * u[5] : T(0), q6 = h6 ? T(0.25) * u[6] : T(0);
        const T e1 = h1 ? T(0.5) * u[1] : T(0), e2 = h2 ? T(0.5) * u[2] : T(0), e4 = h4 ? T(0.5) * u[4] : T(0);
        u[0] = u[0] + (e1 + e2 + e4) + (q3 + q5 + q6);
        u[1] = (h1 ? e1 : u[1]) + q3 + q5;
        u[2] = (h2 ? e2 : u[2]) + q3 + q6;
        u[4] = (h4 ? e4 : u[4]) + q5 + q6;
        if (h3) u[3] = q3;
        if (h5) u[5] = q5;
        if (h6) u[6] = q6;
    }
}

// One element: gather its 8 vertex values from the LDS x slots (byte offsets
// off[j], child frame), apply the element operator and add r into the LDS
// fixed-point accumulators.  Every lane of the wave calls it (the hanging
// test is a wave vote); lanes with valid == false do nothing.
template <typename T>
__device__ __forceinline__ void accumulate_element(const uint32_t *off, uint32_t cm, bool valid,
                                                   const unsigned char *xb, unsigned char *yb, const T (*lam)[4],
                                                   double qs) {
    const uint32_t mask = cm & 255u;
    const bool wave_hanging = __ballot(valid && mask != 0u) != 0ull;
    if (!valid) return;
    T u[8];
    // x slots are sizeof(T) bytes, accumulators 8 bytes: off[] addresses the latter
    constexpr int XSH = sizeof(T) == 4 ? 1 : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) u[j] = *reinterpret_cast<const T *>(xb + (off[j] >> XSH));
    element_apply<T>(u, cm, wave_hanging, lam);
#pragma unroll
    for (int j = 0; j < 8; ++j)
        atomicAdd(reinterpret_cast<unsigned long long *>(yb + off[j]), to_fixed((double)u[j], qs));
}

// The element record holds 8-byte slot offsets; SH = 1 turns them into
// 4-byte slot offsets (FP32 slots).
template <int SH>
__device__ __forceinline__ void unpack_record(const uint4 v, uint32_t *off) {
    off[0] = (v.x & 0xffffu) >> SH;
    off[1] = (v.x >> 16) >> SH;
    off[2] = (v.y & 0xffffu) >> SH;
    off[3] = (v.y >> 16) >> SH;
    off[4] = (v.z & 0xffffu) >> SH;
    off[5] = (v.z >> 16) >> SH;
    off[6] = (v.w & 0xffffu) >> SH;
    off[7] = (v.w >> 16) >> SH;
}
// 12-byte records (DaDevice::lrec): 10-bit local indices of the 8 vertices in
// bits 10j..10j+9 of the 96-bit record, cmeta in bits 80..95.  Offsets come
// out as byte offsets of 8-byte slots, as unpack_record<0> gives them.
__device__ __forceinline__ void unpack_rec12(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t *off) {
    off[0] = (w0 & 1023u) << 3;
    off[1] = ((w0 >> 10) & 1023u) << 3;
    off[2] = ((w0 >> 20) & 1023u) << 3;
    off[3] = (__builtin_amdgcn_alignbit(w1, w0, 30) & 1023u) << 3;
    off[4] = ((w1 >> 8) & 1023u) << 3;
    off[5] = ((w1 >> 18) & 1023u) << 3;
    off[6] = (__builtin_amdgcn_alignbit(w2, w1, 28) & 1023u) << 3;
    off[7] = ((w2 >> 6) & 1023u) << 3;
}
template <int SH>
__device__ __forceinline__ void unpack_offsets(const uint16_t *lconn, size_t e, uint32_t *off) {
    unpack_record<SH>(*reinterpret_cast<const uint4 *>(lconn + e * 8), off);
}

// Pass 1, one workgroup per element block (any block size up to MAXN nodes).
// Element data are loaded first, the block's nodes are gathered into LDS,
// element contributions are accumulated there as fixed-point integers
// (ds_add_u64, exact and order-independent), and each node sum goes to y
// (node touched by this block only) or to its partial slot (shared node).
// Node entries: bits 0-29 node id, bit 30 Dirichlet boundary, bit 31 shared.
// With DIR, boundary columns are dropped and boundary rows are the identity.
// LDS: x slots (8 bytes each) then the accumulators, both static, so the
// per-vertex byte offsets read from lconn address them directly.  Elements
// without hanging vertices come first in each block, so most waves skip the
// hanging interpolation.
template <typename T, bool DIR, int MAXN>
__global__ void __launch_bounds__(B) k_matvec_blk(const uint16_t *__restrict__ lconn, const uint16_t *__restrict__ cmeta,
                                                  size_t ne, const uint32_t *__restrict__ boff,
                                                  const uint32_t *__restrict__ bnode, const uint32_t *__restrict__ bdst,
                                                  const T *__restrict__ x, T *__restrict__ y, T *__restrict__ partial,
                                                  double lap_fac, double mass_fac) {
    constexpr int EB = PER * B;
    __shared__ __align__(16) T xs[MAXN];                   // x of the block's nodes
    __shared__ __align__(16) unsigned long long ys[MAXN];  // fixed-point node sums
    __shared__ T lam[32][4];
    __shared__ int se_tab[32];
    __shared__ int wred[B / 64][2];
    unsigned char *xb = reinterpret_cast<unsigned char *>(xs);
    unsigned char *yb = reinterpret_cast<unsigned char *>(ys);
    const uint32_t b = blockIdx.x, n0 = boff[b], nb = boff[b + 1] - n0;
    const size_t e0 = (size_t)b * EB;
    uint32_t off[PER][8], mt[PER];
    int lv = 31;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const size_t e = e0 + threadIdx.x + k * B;
        mt[k] = 0;
        if (e < ne) {
            unpack_offsets<0>(lconn, e, off[k]);
            mt[k] = cmeta[e];
            lv = min(lv, (int)(mt[k] >> 8));
        }
    }
    init_lambda<T>(lam, lap_fac, mass_fac);
    init_level_exps(se_tab, lap_fac, mass_fac);
    int xe = kExpZero;
    for (uint32_t l = threadIdx.x; l < nb; l += B) {
        const uint32_t g = bnode[n0 + l];
        const T v = (DIR && (g & BDY_BIT)) ? T(0) : x[g & ID_MASK];
        xs[l] = v;
        ys[l] = 0ull;
        xe = max(xe, mag_exp((double)v));
    }
    __syncthreads();  // se_tab
    publish_scale_inputs(xe, se_tab[lv], wred);
    __syncthreads();
    const FixScale fs = block_scale(wred);
#pragma unroll
    for (int k = 0; k < PER; ++k)
        accumulate_element<T>(off[k], mt[k], e0 + threadIdx.x + k * B < ne, xb, yb, lam, fs.s);
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < nb; l += B) {
        const uint32_t g = bnode[n0 + l];
        const T v = (T)from_fixed(ys[l], fs);
        if (g & SHARED_BIT) partial[bdst[n0 + l]] = v;
        else y[g & ID_MASK] = (DIR && (g & BDY_BIT)) ? x[g & ID_MASK] : v;
    }
}

// Pass 1, persistent and software-pipelined, for blocks of at most NPT * B
// nodes.  The grid is a multiple of 8 and workgroup w runs on XCD w % 8; each
// XCD walks a contiguous range of element blocks, so neighbouring blocks
// (which share nodes and x cache lines) meet in the same L2.  A workgroup
// works on block b while the loads of the blocks after it are in flight:
// the x values, element records and partial slots of b+1, the node entries of
// b+2 and the node offsets of b+3 (b+1 = the workgroup's next block).  No load in the
// loop depends on a load issued in the same iteration, so the element phase
// never waits for memory.  Loads past a block's node count are clamped to its
// last entry and element records are read for whole blocks (the record
// buffers are padded to whole blocks); the values loaded there are not used.
// EPI selects what the writeout stores for a node with row value v = (A x)_i:
// EPI_Y: y_i = v; EPI_JACOBI: x_i += omega dinv_i (b_i - v), in place (y is
// x); EPI_RESID: y_i = b_i - v.  In place is safe: a block writes only nodes
// no other block reads, and shared nodes are finished by the second pass.
template <typename T, bool DIR, int NPT, int EPI>
__global__ void __launch_bounds__(B, 4) k_matvec_pipe(const uint32_t *__restrict__ rec, size_t ne, uint32_t nblk,
                                                   const uint32_t *__restrict__ boff,
                                                   const uint32_t *__restrict__ bnode,
                                                   const uint32_t *__restrict__ bdst, const T *x, T *y,
                                                   T *__restrict__ partial, double lap_fac, double mass_fac,
                                                   const T *__restrict__ bvec, const T *__restrict__ dinv,
                                                   T omega) {
    constexpr int EB = PER * B, MAXN = NPT * B;
    __shared__ __align__(16) T xs[MAXN];
    __shared__ __align__(16) unsigned long long ys[MAXN];  // fixed-point node sums
    __shared__ int se_tab[32];
    __shared__ int wred[B / 64][2];
    __shared__ T lam[32][4];
    unsigned char *xb = reinterpret_cast<unsigned char *>(xs);
    unsigned char *yb = reinterpret_cast<unsigned char *>(ys);
    const uint32_t xcd = blockIdx.x & 7u, step = gridDim.x >> 3;
    const uint32_t lo = (uint32_t)((uint64_t)xcd * nblk / 8), hi = (uint32_t)((uint64_t)(xcd + 1) * nblk / 8);
    uint32_t b = lo + (blockIdx.x >> 3);
    if (b >= hi) return;  // uniform over the workgroup
    init_lambda<T>(lam, lap_fac, mass_fac);
    init_level_exps(se_tab, lap_fac, mass_fac);

    auto load_elems = [&](uint32_t bb, uint3 *ev) {
        const size_t e = (size_t)bb * EB + threadIdx.x;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t *r = rec + (e + k * B) * 3;
            ev[k].x = r[0];
            ev[k].y = r[1];
            ev[k].z = r[2];
        }
    };
    auto load_entries = [&](uint32_t s0, uint32_t n, uint32_t *g) {
#pragma unroll
        for (int i = 0; i < NPT; ++i) g[i] = bnode[s0 + min(threadIdx.x + i * B, n - 1u)];
    };
    auto gather = [&](const uint32_t *g, T *v) {
#pragma unroll
        for (int i = 0; i < NPT; ++i) v[i] = x[g[i] & ID_MASK];
    };
    // partial slots (read for every entry, used by the shared ones)
    auto load_slots = [&](uint32_t s0, uint32_t n, uint32_t *d) {
#pragma unroll
        for (int i = 0; i < NPT; ++i) d[i] = bdst[s0 + min(threadIdx.x + i * B, n - 1u)];
    };

    // block b: element records, node entries, x values
    uint32_t s0 = boff[b], nb = boff[b + 1] - s0;
    uint3 ev[PER];
    uint32_t g[NPT], sd[NPT];
    T xv[NPT];
    load_elems(b, ev);
    load_entries(s0, nb, g);
    load_slots(s0, nb, sd);
    gather(g, xv);
    // block bA = b + step: node entries; block bB: node offsets
    uint32_t bA = b + step, s0A = 0, nA = 1, gA[NPT];
    if (bA < hi) {
        s0A = boff[bA];
        nA = boff[bA + 1] - s0A;
        load_entries(s0A, nA, gA);
    }
    uint32_t bB = bA + step, s0B = 0, nB = 1;
    if (bB < hi) {
        s0B = boff[bB];
        nB = boff[bB + 1] - s0B;
    }
    // Drain the prologue loads here: the loop body's wait counts then only
    // see the loads issued inside the loop, and the element phase does not
    // wait for the loads issued ahead of it in the same iteration.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // lam, se_tab
    for (;;) {
        // x of the block's nodes for the identity rows of boundary nodes (the
        // Jacobi writeout reads x back from the LDS slots instead and reloads
        // it for boundary nodes, whose slots hold 0)
        T xr[NPT];
        int xe = kExpZero, lv = 31;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const uint32_t l = threadIdx.x + i * B;
            if (l < nb) {
                const T v = (DIR && (g[i] & BDY_BIT)) ? T(0) : xv[i];
                xs[l] = v;
                ys[l] = 0ull;
                xe = max(xe, mag_exp((double)v));
            }
            if (DIR && EPI != EPI_JACOBI) xr[i] = xv[i];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if ((size_t)b * EB + threadIdx.x + k * B < ne) lv = min(lv, (int)(ev[k].z >> 24));
        publish_scale_inputs(xe, se_tab[lv], wred);
        // right-hand side and inverse diagonal of the block's nodes, for the writeout
        T bv[NPT], dv[NPT];
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            if (EPI != EPI_Y) bv[i] = bvec[g[i] & ID_MASK];
            if (EPI == EPI_JACOBI) dv[i] = dinv[g[i] & ID_MASK];
        }
        __syncthreads();
        const FixScale fs = block_scale(wred);
        const bool moreA = bA < hi;
        uint3 evA[PER];
        uint32_t gB[NPT], sdA[NPT];
        if (moreA) {
            gather(gA, xv);
            load_elems(bA, evA);
            load_slots(s0A, nA, sdA);
        }
        if (bB < hi) load_entries(s0B, nB, gB);
        const uint32_t bC = bB + step;
        uint32_t s0C = 0, nC = 1;
        if (bC < hi) {
            s0C = boff[bC];
            nC = boff[bC + 1] - s0C;
        }
        const size_t e0 = (size_t)b * EB + threadIdx.x;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            uint32_t off[8];
            unpack_rec12(ev[k].x, ev[k].y, ev[k].z, off);
            accumulate_element<T>(off, ev[k].z >> 16, e0 + k * B < ne, xb, yb, lam, fs.s);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            const uint32_t l = threadIdx.x + i * B;
            if (l < nb) {
                const T av = (T)from_fixed(ys[l], fs);
                if (g[i] & SHARED_BIT) {
                    partial[sd[i]] = av;
                } else if (EPI == EPI_JACOBI) {
                    const T xi = (DIR && (g[i] & BDY_BIT)) ? x[g[i] & ID_MASK] : xs[l];
                    const T v = (DIR && (g[i] & BDY_BIT)) ? xi : av;
                    y[g[i] & ID_MASK] = xi + omega * dv[i] * (bv[i] - v);
                } else {
                    const T v = (DIR && (g[i] & BDY_BIT)) ? xr[i] : av;
                    if (EPI == EPI_Y) y[g[i] & ID_MASK] = v;
                    else y[g[i] & ID_MASK] = bv[i] - v;
                }
            }
        }
        if (!moreA) break;
        __syncthreads();  // the LDS of block b is reused
        b = bA;
        s0 = s0A;
        nb = nA;
#pragma unroll
        for (int k = 0; k < PER; ++k) ev[k] = evA[k];
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
            g[i] = gA[i];
            sd[i] = sdA[i];
            gA[i] = gB[i];
        }
        bA = bB;
        s0A = s0B;
        nA = nB;
        bB = bC;
        s0B = s0C;
        nB = nC;
    }
}

// Pass 2: shared nodes sum their partials in block order and store as the
// first pass does (EPI).
template <typename T, bool DIR, int EPI>
__global__ void __launch_bounds__(B) k_matvec_shared(const uint32_t *__restrict__ soff,
                                                     const uint32_t *__restrict__ snode, uint32_t ns,
                                                     const T *__restrict__ partial, const T *x, T *y,
                                                     const T *__restrict__ bvec, const T *__restrict__ dinv,
                                                     T omega) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i >= ns) return;
    const uint32_t g = snode[i], id = g & ID_MASK;
    // the first 8 partials and the writeout's operands are loaded together;
    // the sum runs in slot (block) order
    const uint32_t p0 = soff[i], p1 = soff[i + 1];
    T pv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) pv[k] = p0 + k < p1 ? partial[p0 + k] : T(0);
    const T xi = (DIR || EPI == EPI_JACOBI) ? x[id] : T(0);
    const T bi = EPI != EPI_Y ? bvec[id] : T(0);
    const T di = EPI == EPI_JACOBI ? dinv[id] : T(0);
    T v = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (p0 + k < p1) v += pv[k];
    for (uint32_t p = p0 + 8; p < p1; ++p) v += partial[p];
    if (DIR && (g & BDY_BIT)) v = xi;
    if (EPI == EPI_Y) y[id] = v;
    else if (EPI == EPI_RESID) y[id] = bi - v;
    else y[id] = xi + omega * di * (bi - v);
}
// Stand-alone writeouts for the blocks the pipelined kernel does not take.
template <typename T>
__global__ void k_after_jacobi(const T *bvec, const T *ax, const T *dinv, T omega, uint32_t n, T *x) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) x[i] = x[i] + omega * dinv[i] * (bvec[i] - ax[i]);
}
template <typename T>
__global__ void k_after_resid(const T *bvec, uint32_t n, T *y) {
    uint32_t i = blockIdx.x * B + threadIdx.x;
    if (i < n) y[i] = bvec[i] - y[i];
}

// Persistent grid: resident workgroups per CU x CUs, a multiple of 8 (one
// share per XCD), capped so that each workgroup gets at least one block.
unsigned pipe_grid(const void *kernel, uint32_t nblk) {
    static thread_local int cached_dev = -1, cus = 0;
    static thread_local std::vector<std::pair<const void *, int>> occ;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev != cached_dev) {
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        occ.clear();
        cached_dev = dev;
    }
    int per_cu = 0;
    for (auto &kv : occ)
        if (kv.first == kernel) per_cu = kv.second;
    if (!per_cu) {
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, B, 0));
        if (per_cu < 1) per_cu = 1;
        occ.emplace_back(kernel, per_cu);
    }
    uint64_t g = (uint64_t)per_cu * (uint64_t)cus;
    const uint64_t need = ((uint64_t)nblk + 7) / 8 * 8;
    if (g > need) g = need;
    g = (g + 7) / 8 * 8;
    return (unsigned)std::max<uint64_t>(g, 8);
}

// EPI as for k_matvec_pipe; with EPI_JACOBI, y is x and scratch (n_nodes
// entries) holds A x on the path for DAs whose blocks exceed the pipelined
// kernel's 1024 nodes.
// partial_at: where the shared nodes' partial sums go (nullptr: the DA's own
// buffer of partial_slots entries).
template <typename T, bool DIR, int EPI>
void matvec_blocks(const DaDevice &da, const T *x, T *y, T lap_fac, T mass_fac, const T *bvec, const T *dinv,
                   T omega, T *scratch, hipStream_t s, T *partial_at = nullptr) {
    if (!da.n_elem) return;
    if (da.elem_block != PER * B) {
        dgx_set_error("matvec: DA built with element blocks of %u, kernel expects %d", da.elem_block, PER * B);
        throw DgxError();
    }
    T *partial = partial_at ? partial_at : reinterpret_cast<T *>(da.partial.p);
    const uint32_t mn = da.max_block_nodes;
    // DGX_MATVEC_BLOCKED=1 takes the path of DAs with large blocks for every
    // DA (a test hook: real meshes rarely reach it)
    const char *forced = getenv("DGX_MATVEC_BLOCKED");
    static_assert(4 * B == kPipeMaxNodes, "pipelined matvec block size");
    if (mn <= kPipeMaxNodes && da.lrec.p && !(forced && forced[0] == '1')) {
        const unsigned grid =
            pipe_grid(reinterpret_cast<const void *>(&k_matvec_pipe<T, DIR, 4, EPI>), da.n_blocks);
        k_matvec_pipe<T, DIR, 4, EPI><<<grid, B, 0, s>>>(da.lrec.p, da.n_elem, da.n_blocks,
                                                         da.bnode_off.p, da.bnode.p, da.bdst.p, x, y, partial,
                                                         (double)lap_fac, (double)mass_fac, bvec, dinv, omega);
        if (da.n_shared)
            k_matvec_shared<T, DIR, EPI><<<grid_for(da.n_shared, B), B, 0, s>>>(
                da.soff.p, da.snode.p, da.n_shared, partial, x, y, bvec, dinv, omega);
        HIP_OK(hipGetLastError());
        return;
    }
    T *ax = EPI == EPI_JACOBI ? scratch : y;
#define DGX_MV_LAUNCH(MAXN)                                                                                     \
    k_matvec_blk<T, DIR, MAXN><<<da.n_blocks, B, 0, s>>>(da.lconn.p, da.cmeta.p, da.n_elem, da.bnode_off.p,     \
                                                          da.bnode.p, da.bdst.p, x, ax, partial, (double)lap_fac, \
                                                          (double)mass_fac)
    if (mn <= 2048) DGX_MV_LAUNCH(2048);
    else if (mn <= 4096) DGX_MV_LAUNCH(4096);
    else {
        dgx_set_error("matvec: block with %u nodes exceeds 4096", mn);
        throw DgxError();
    }
#undef DGX_MV_LAUNCH
    if (da.n_shared)
        k_matvec_shared<T, DIR, EPI_Y><<<grid_for(da.n_shared, B), B, 0, s>>>(da.soff.p, da.snode.p, da.n_shared,
                                                                              partial, x, ax, nullptr, nullptr, T(0));
    const unsigned g = grid_for(da.n_nodes, B);
    if (EPI == EPI_RESID) k_after_resid<T><<<g, B, 0, s>>>(bvec, da.n_nodes, y);
    if (EPI == EPI_JACOBI) k_after_jacobi<T><<<g, B, 0, s>>>(bvec, ax, dinv, omega, da.n_nodes, y);
    HIP_OK(hipGetLastError());
}

}  // namespace

void da_matvec_f64(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac, hipStream_t s) {
    matvec_blocks<double, false, EPI_Y>(da, x, y, lap_fac, mass_fac, nullptr, nullptr, 0.0, nullptr, s);
}
// ---------------------------------------------------------------------------
// Matvec without allocation: the shared nodes' partial sums go to a static
// device array instead of the DA's buffer.  Each device holds its own copy of
// the array; host code takes its address from hipGetSymbolAddress once per
// device and passes that pointer to the kernels (the array's name in host code
// is not a device address).
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchDoubles = kNoAllocScratchBytes / sizeof(double);
__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

double *scratch_ptr_of_current_device() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        dgx_set_error("matvec (no allocation): device %d out of range", dev);
        throw DgxError();
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            dgx_set_error("matvec (no allocation): hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            throw DgxError();
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void da_noalloc_init() { (void)scratch_ptr_of_current_device(); }

void da_matvec_f64_noalloc(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac,
                           hipStream_t s) {
    if (da.partial_slots > kNoAllocScratchDoubles) {
        dgx_set_error("matvec (no allocation): the DA has %llu shared-node partial slots, above the %zu that the "
                      "static scratch array of %zu bytes holds",
                      (unsigned long long)da.partial_slots, kNoAllocScratchDoubles, kNoAllocScratchBytes);
        throw DgxError();
    }
    matvec_blocks<double, false, EPI_Y>(da, x, y, lap_fac, mass_fac, nullptr, nullptr, 0.0, nullptr, s,
                                        scratch_ptr_of_current_device());
}

void da_matvec_f32(const DaDevice &da, const float *x, float *y, float lap_fac, float mass_fac, hipStream_t s) {
    matvec_blocks<float, false, EPI_Y>(da, x, y, lap_fac, mass_fac, nullptr, nullptr, 0.0f, nullptr, s);
}
void da_matvec_dirichlet_f64(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac,
                             hipStream_t s) {
    matvec_blocks<double, true, EPI_Y>(da, x, y, lap_fac, mass_fac, nullptr, nullptr, 0.0, nullptr, s);
}
void da_matvec_dirichlet_f32(const DaDevice &da, const float *x, float *y, float lap_fac, float mass_fac,
                             hipStream_t s) {
    matvec_blocks<float, true, EPI_Y>(da, x, y, lap_fac, mass_fac, nullptr, nullptr, 0.0f, nullptr, s);
}
template <typename T>
void da_dirichlet_jacobi(const DaDevice &da, T *x, const T *b, const T *dinv, T omega, T lap_fac, T mass_fac,
                         T *scratch, hipStream_t s) {
    matvec_blocks<T, true, EPI_JACOBI>(da, x, x, lap_fac, mass_fac, b, dinv, omega, scratch, s);
}
template <typename T>
void da_dirichlet_residual(const DaDevice &da, const T *x, const T *b, T *r, T lap_fac, T mass_fac,
                           hipStream_t s) {
    matvec_blocks<T, true, EPI_RESID>(da, x, r, lap_fac, mass_fac, b, nullptr, T(0), nullptr, s);
}
template void da_dirichlet_jacobi<double>(const DaDevice &, double *, const double *, const double *, double, double,
                                          double, double *, hipStream_t);
template void da_dirichlet_jacobi<float>(const DaDevice &, float *, const float *, const float *, float, float, float,
                                         float *, hipStream_t);
template void da_dirichlet_residual<double>(const DaDevice &, const double *, const double *, double *, double, double,
                                            hipStream_t);
template void da_dirichlet_residual<float>(const DaDevice &, const float *, const float *, float *, float, float,
                                           hipStream_t);

namespace {

// Diagonal (C^T A_e C)[m][m] of each element in its child frame: the element
// operator applied to the unit vector of vertex m, entry m.  Summed per block
// with the matvec's fixed-point accumulators (|x| = 1 in the bound), shared
// nodes through the partial slots, so the diagonal is deterministic too.
template <int MAXN>
__global__ void __launch_bounds__(B) k_diag_blk(const uint16_t *__restrict__ lconn, const uint16_t *__restrict__ cmeta,
                                                size_t ne, const uint32_t *__restrict__ boff,
                                                const uint32_t *__restrict__ bnode, const uint32_t *__restrict__ bdst,
                                                double *__restrict__ diag, double *__restrict__ partial,
                                                double lap_fac, double mass_fac) {
    constexpr int EB = PER * B;
    __shared__ __align__(16) unsigned long long ys[MAXN];
    __shared__ double lam[32][4];
    __shared__ int se_tab[32];
    __shared__ int wred[B / 64][2];
    unsigned char *yb = reinterpret_cast<unsigned char *>(ys);
    const uint32_t b = blockIdx.x, n0 = boff[b], nb = boff[b + 1] - n0;
    const size_t e0 = (size_t)b * EB;
    uint32_t off[PER][8], mt[PER];
    int lv = 31;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const size_t e = e0 + threadIdx.x + k * B;
        mt[k] = 0;
        if (e < ne) {
            unpack_offsets<0>(lconn, e, off[k]);
            mt[k] = cmeta[e];
            lv = min(lv, (int)(mt[k] >> 8));
        }
    }
    init_lambda<double>(lam, lap_fac, mass_fac);
    init_level_exps(se_tab, lap_fac, mass_fac);
    for (uint32_t l = threadIdx.x; l < nb; l += B) ys[l] = 0ull;
    __syncthreads();  // se_tab
    publish_scale_inputs(mag_exp(1.0), se_tab[lv], wred);
    __syncthreads();
    const FixScale fs = block_scale(wred);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (e0 + threadIdx.x + k * B >= ne) continue;
        for (int m = 0; m < 8; ++m) {
            double u[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) u[j] = j == m ? 1.0 : 0.0;
            element_apply<double>(u, mt[k], (mt[k] & 255u) != 0u, lam);
            double dm = u[0];
#pragma unroll
            for (int j = 1; j < 8; ++j)
                if (j == m) dm = u[j];
            atomicAdd(reinterpret_cast<unsigned long long *>(yb + off[k][m]), to_fixed(dm, fs.s));
        }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < nb; l += B) {
        const uint32_t g = bnode[n0 + l];
        const double v = from_fixed(ys[l], fs);
        if (g & SHARED_BIT) partial[bdst[n0 + l]] = v;
        else diag[g & ID_MASK] = v;
    }
}

}  // namespace

void da_diagonal_f64(const DaDevice &da, double *diag, double lap_fac, double mass_fac, hipStream_t s) {
    if (!da.n_elem) {
        HIP_OK(hipMemsetAsync(diag, 0, (size_t)da.n_nodes * sizeof(double), s));
        return;
    }
    if (da.elem_block != PER * B) {
        dgx_set_error("diagonal: DA built with element blocks of %u, kernel expects %d", da.elem_block, PER * B);
        throw DgxError();
    }
    double *partial = reinterpret_cast<double *>(da.partial.p);
    const uint32_t mn = da.max_block_nodes;
#define DGX_DIAG_LAUNCH(MAXN)                                                                                    \
    k_diag_blk<MAXN><<<da.n_blocks, B, 0, s>>>(da.lconn.p, da.cmeta.p, da.n_elem, da.bnode_off.p, da.bnode.p,  \
                                               da.bdst.p, diag, partial, lap_fac, mass_fac)
    if (mn <= 1024) DGX_DIAG_LAUNCH(1024);
    else if (mn <= 2048) DGX_DIAG_LAUNCH(2048);
    else if (mn <= 4096) DGX_DIAG_LAUNCH(4096);
    else {
        dgx_set_error("diagonal: block with %u nodes exceeds 4096", mn);
        throw DgxError();
    }
#undef DGX_DIAG_LAUNCH
    if (da.n_shared)
        k_matvec_shared<double, false, EPI_Y><<<grid_for(da.n_shared, B), B, 0, s>>>(
            da.soff.p, da.snode.p, da.n_shared, partial, nullptr, diag, nullptr, nullptr, 0.0);
    HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Element matrices (coarsest-level direct solver).
// ---------------------------------------------------------------------------
namespace {
// vals[e*64 + k*8 + j] = entry (row j, column k) of element e's matrix in its
// child frame: the contribution to A[id[j^c], id[k^c]] (c the child number).
__global__ void __launch_bounds__(B) k_element_matrices(const uint16_t *__restrict__ meta, size_t ne, double lap_fac,
                                                        double mass_fac, double *__restrict__ vals) {
    size_t e = (size_t)blockIdx.x * B + threadIdx.x;
    if (e >= ne) return;
    const uint32_t mt = meta[e], c = (mt >> 8) & 7u, mask = mt & 255u;
    uint32_t cm = 0;
    for (uint32_t j = 0; j < 8; ++j)
        if ((mask >> (j ^ c)) & 1u) cm |= 1u << j;
    cm |= (mt >> 11) << 8;
    for (int k = 0; k < 8; ++k) {
        double xv[8], r[8];
        for (int j = 0; j < 8; ++j) xv[j] = (j == k) ? 1.0 : 0.0;
        element_child_frame<double>(cm, lap_fac, mass_fac, xv, r);
        for (int j = 0; j < 8; ++j) vals[e * 64 + k * 8 + j] = r[j];
    }
}
}  // namespace

void da_element_matrices(const DaDevice &da, double lap_fac, double mass_fac, double *vals, hipStream_t s) {
    if (!da.n_elem) return;
    k_element_matrices<<<grid_for(da.n_elem, B), B, 0, s>>>(da.meta.p, da.n_elem, lap_fac, mass_fac, vals);
    HIP_OK(hipGetLastError());
}
