// dgx_matvec.hip -- element matvec of -lap + I on the octree DA (gfx950).
//
// Reference: Jacobian1MatMult (examples/src/backend/odaJac.C:275-366) applies,
// per element, out[idx_k] += sum_j (h/2 L[c][t][k][j] + h^3/8 M[c][t][k][j]) in[idx_j]
// with the type-2 stencils of handleType2Stencils.C:678-1000.  Those tables
// are C^T K C where K is the Q1 stiffness/mass on [-1,1]^3 and C interpolates
// the element's hanging vertices from its parent's corners: a vertex hanging on
// a parent edge takes the mean of the two edge ends, one hanging on a parent
// face the mean of the four face corners (checked against the reference tables
// to 8e-16 by the test suite).  The kernel applies C, K and C^T directly from
// the 8-bit hanging mask and the child number, so no stencil table is read:
// per element it moves 32 B of node ids + 2 B of meta + 8 gathers + 8 scatters.
//
// K is applied by sum factorisation with the 1-D operators
//   m = [[2/3,1/3],[1/3,2/3]] (mass), k = [[1/2,-1/2],[-1/2,1/2]] (stiffness):
//   K u = (k(x) m(y) m(z) + m(x) k(y) m(z) + m(x) m(y) k(z)) u,  M u = m(x) m(y) m(z) u.
#include "dgx_da.h"

namespace {

constexpr int B = 256;

template <typename T>
struct Elem8 {
    T v[8];
};

// 1-D operators along axis a (stride s = 1, 2, 4) on an 8-vector.
template <typename T, int S>
__device__ __forceinline__ void op_m(const T *u, T *o) {
    const T third = T(1.0 / 3.0);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & S) continue;
        T a = u[i], b = u[i + S];
        o[i] = (a + a + b) * third;
        o[i + S] = (a + b + b) * third;
    }
}
template <typename T, int S>
__device__ __forceinline__ void op_k(const T *u, T *o) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i & S) continue;
        T d = T(0.5) * (u[i] - u[i + S]);
        o[i] = d;
        o[i + S] = -d;
    }
}

// v = fac1 * K u + fac2 * M u
template <typename T>
__device__ __forceinline__ void apply_q1(const T *u, T fac1, T fac2, T *v) {
    T a[8], b[8], c[8], d[8], e[8], w[8], kb[8];
    op_m<T, 4>(u, a);   // Mz u
    op_m<T, 2>(a, b);   // My Mz u
    op_k<T, 2>(a, c);   // Ky Mz u
    op_k<T, 4>(u, d);   // Kz u
    op_m<T, 2>(d, e);   // My Kz u
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = fac1 * (c[i] + e[i]) + fac2 * b[i];
    op_m<T, 1>(w, v);   // Mx (...)
    op_k<T, 1>(b, kb);  // Kx My Mz u
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] += fac1 * kb[i];
}

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

template <typename T>
__global__ void __launch_bounds__(B) k_matvec(const uint32_t *__restrict__ conn, const uint16_t *__restrict__ meta,
                                              size_t ne, const T *__restrict__ x, T *__restrict__ y, T lap_fac,
                                              T mass_fac) {
    size_t e = (size_t)blockIdx.x * B + threadIdx.x;
    if (e >= ne) return;
    uint32_t mt = meta[e];
    uint32_t mask = mt & 255u, c = (mt >> 8) & 7u, lev = mt >> 11;
    T h = T(1.0) / T(1u << lev);
    T fac1 = lap_fac * h * T(0.5), fac2 = mass_fac * h * h * h * T(0.125);
    uint32_t idx[8];
    T xv[8], u[8], v[8], r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) idx[j] = conn[(size_t)j * ne + e];
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = x[idx[j]];
    if (mask == 0) {
        apply_q1<T>(xv, fac1, fac2, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) atomic_add(&y[idx[j]], v[j]);
        return;
    }
    // u = C x: hanging vertices interpolate from the parent's corners
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        if (!((mask >> a) & 1u)) {
            u[a] = xv[a];
            continue;
        }
        uint32_t dff = a ^ c, lo = dff & (0u - dff);
        if (__popc(dff) == 1) {
            u[a] = T(0.5) * (xv[c] + xv[a]);
        } else {
            u[a] = T(0.25) * (xv[c] + xv[a] + xv[c ^ lo] + xv[c ^ (dff ^ lo)]);
        }
    }
    apply_q1<T>(u, fac1, fac2, v);
    // r = C^T v
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = T(0);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        if (!((mask >> a) & 1u)) {
            r[a] += v[a];
            continue;
        }
        uint32_t dff = a ^ c, lo = dff & (0u - dff);
        if (__popc(dff) == 1) {
            T hv = T(0.5) * v[a];
            r[c] += hv;
            r[a] += hv;
        } else {
            T qv = T(0.25) * v[a];
            r[c] += qv;
            r[a] += qv;
            r[c ^ lo] += qv;
            r[c ^ (dff ^ lo)] += qv;
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) atomic_add(&y[idx[j]], r[j]);
}

}  // namespace

void da_matvec_f64(const DaDevice &da, const double *x, double *y, double lap_fac, double mass_fac, hipStream_t s) {
    HIP_OK(hipMemsetAsync(y, 0, (size_t)da.n_nodes * sizeof(double), s));
    if (!da.n_elem) return;
    k_matvec<double><<<grid_for(da.n_elem, B), B, 0, s>>>(da.conn.p, da.meta.p, da.n_elem, x, y, lap_fac, mass_fac);
    HIP_OK(hipGetLastError());
}

void da_matvec_f32(const DaDevice &da, const float *x, float *y, float lap_fac, float mass_fac, hipStream_t s) {
    HIP_OK(hipMemsetAsync(y, 0, (size_t)da.n_nodes * sizeof(float), s));
    if (!da.n_elem) return;
    k_matvec<float><<<grid_for(da.n_elem, B), B, 0, s>>>(da.conn.p, da.meta.p, da.n_elem, x, y, lap_fac, mass_fac);
    HIP_OK(hipGetLastError());
}

namespace {

// diag[idx_k] += (C^T K C)[k][k]: column k of C is the interpolation weight of
// DOF k at each vertex; the diagonal entry is that column's K-norm.
__global__ void __launch_bounds__(B) k_diagonal(const uint32_t *__restrict__ conn, const uint16_t *__restrict__ meta,
                                                size_t ne, double *__restrict__ diag, double lap_fac,
                                                double mass_fac) {
    size_t e = (size_t)blockIdx.x * B + threadIdx.x;
    if (e >= ne) return;
    uint32_t mt = meta[e];
    uint32_t mask = mt & 255u, c = (mt >> 8) & 7u, lev = mt >> 11;
    double h = 1.0 / (double)(1u << lev);
    double fac1 = lap_fac * h * 0.5, fac2 = mass_fac * h * h * h * 0.125;
    for (int k = 0; k < 8; ++k) {
        double u[8], v[8];
        for (int a = 0; a < 8; ++a) {
            double w = 0.0;
            if (!((mask >> a) & 1u)) {
                w = (a == k) ? 1.0 : 0.0;
            } else {
                uint32_t dff = a ^ c, lo = dff & (0u - dff);
                if (__popc(dff) == 1) {
                    w = ((uint32_t)k == c || k == a) ? 0.5 : 0.0;
                } else {
                    w = ((uint32_t)k == c || k == a || (uint32_t)k == (c ^ lo) || (uint32_t)k == (c ^ (dff ^ lo))) ? 0.25
                                                                                                                  : 0.0;
                }
            }
            u[a] = w;
        }
        apply_q1<double>(u, fac1, fac2, v);
        double dk = 0.0;
        for (int a = 0; a < 8; ++a) dk += u[a] * v[a];
        atomic_add(&diag[conn[(size_t)k * ne + e]], dk);
    }
}

}  // namespace

void da_diagonal_f64(const DaDevice &da, double *diag, double lap_fac, double mass_fac, hipStream_t s) {
    HIP_OK(hipMemsetAsync(diag, 0, (size_t)da.n_nodes * sizeof(double), s));
    if (!da.n_elem) return;
    k_diagonal<<<grid_for(da.n_elem, B), B, 0, s>>>(da.conn.p, da.meta.p, da.n_elem, diag, lap_fac, mass_fac);
    HIP_OK(hipGetLastError());
}
