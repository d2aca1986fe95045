// gfx950 kernels for tet4 and hex8 meshes: one element per lane.
//
// Device layout: idx int32 SoA [nv][E] and geometry SoA [k][E] (the layout of the
// reference's device path, operators/cuda/cu_laplacian.hpp fff_stride), so a wave
// reads 64 consecutive words of every row: fully coalesced.  Scatter with FP64
// atomics, as cu_tet4_laplacian_apply_kernel (operators/tet4/cuda/cu_tet4_laplacian.cu:11-44).
#include "common.hpp"
#include "unstructured_kernels.hpp"

namespace sfem_amd {

// ------------------------------------------------------------------ tet4 Laplacian
// e = K u with K from FFF (operators/tet4/tet4_laplacian_inline_cpu.hpp:54-73):
// grad u = (u1-u0, u2-u0, u3-u0) on the reference tet, t = FFF grad u,
// e1..3 = t, e0 = -(t0+t1+t2).
__global__ void tet4_laplacian_apply_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                            const float *__restrict__ fff, const double *__restrict__ x,
                                            double *__restrict__ y) {
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        int32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = idx[k * E + e];
        double f[6];
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = (double)fff[k * E + e];
        const double u0 = x[v[0]];
        const double g0 = x[v[1]] - u0, g1 = x[v[2]] - u0, g2 = x[v[3]] - u0;
        const double t0 = f[0] * g0 + f[1] * g1 + f[2] * g2;
        const double t1 = f[1] * g0 + f[3] * g1 + f[4] * g2;
        const double t2 = f[2] * g0 + f[4] * g1 + f[5] * g2;
        atomicAdd(&y[v[0]], -(t0 + t1 + t2));
        atomicAdd(&y[v[1]], t0);
        atomicAdd(&y[v[2]], t1);
        atomicAdd(&y[v[3]], t2);
    }
}

__global__ void tet4_laplacian_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                           const float *__restrict__ fff, double *__restrict__ d) {
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        double f[6];
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = (double)fff[k * E + e];
        // tet4_laplacian_diag_fff (operators/tet4/tet4_laplacian_inline_cpu.hpp:32-41)
        atomicAdd(&d[idx[0 * E + e]], f[0] + 2 * f[1] + 2 * f[2] + f[3] + 2 * f[4] + f[5]);
        atomicAdd(&d[idx[1 * E + e]], f[0]);
        atomicAdd(&d[idx[2 * E + e]], f[3]);
        atomicAdd(&d[idx[3 * E + e]], f[5]);
    }
}

// ------------------------------------------------------------------ hex8 Laplacian
// Exact Q1 integral with constant FFF by 2-point Gauss (the integrand is at most
// quadratic per direction): y_i = sum_q w_q grad phi_i(q) . G grad u(q).
__device__ __forceinline__ void q1_ref_grad(int n, double qx, double qy, double qz, double *g) {
    const int bx = hex8_cx(n), by = hex8_cy(n), bz = hex8_cz(n);
    const double fx = bx ? qx : 1 - qx, fy = by ? qy : 1 - qy, fz = bz ? qz : 1 - qz;
    const double sx = bx ? 1 : -1, sy = by ? 1 : -1, sz = bz ? 1 : -1;
    g[0] = sx * fy * fz;
    g[1] = fx * sy * fz;
    g[2] = fx * fy * sz;
}

__global__ void hex8_laplacian_apply_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                            const float *__restrict__ fff, const float *__restrict__ coeff,
                                            const double *__restrict__ x, double *__restrict__ y) {
    const double gp[2] = {0.5 - 0.28867513459481288225, 0.5 + 0.28867513459481288225};
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        int32_t v[8];
        double u[8], out[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k] = idx[k * E + e];
            u[k] = x[v[k]];
            out[k] = 0;
        }
        const double c = (coeff ? (double)coeff[e] : 1.0) * 0.125;
        double f[6];
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = (double)fff[k * E + e] * c;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double qx = gp[q & 1], qy = gp[(q >> 1) & 1], qz = gp[q >> 2];
            double gu[3] = {0, 0, 0};
#pragma unroll
            for (int n = 0; n < 8; n++) {
                double g[3];
                q1_ref_grad(n, qx, qy, qz, g);
                gu[0] += g[0] * u[n];
                gu[1] += g[1] * u[n];
                gu[2] += g[2] * u[n];
            }
            const double t0 = f[0] * gu[0] + f[1] * gu[1] + f[2] * gu[2];
            const double t1 = f[1] * gu[0] + f[3] * gu[1] + f[4] * gu[2];
            const double t2 = f[2] * gu[0] + f[4] * gu[1] + f[5] * gu[2];
#pragma unroll
            for (int n = 0; n < 8; n++) {
                double g[3];
                q1_ref_grad(n, qx, qy, qz, g);
                out[n] += g[0] * t0 + g[1] * t1 + g[2] * t2;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) atomicAdd(&y[v[k]], out[k]);
    }
}

__global__ void hex8_laplacian_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                           const float *__restrict__ fff, const float *__restrict__ coeff,
                                           double *__restrict__ d) {
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        const double c = coeff ? (double)coeff[e] : 1.0;
        double f[6];
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = (double)fff[k * E + e] * c;
#pragma unroll
        for (int k = 0; k < 8; k++) atomicAdd(&d[idx[k * E + e]], hex8_laplacian_entry(f, k, k));
    }
}

// ------------------------------------------------------------------ hex8 elasticity
// Affine Q1 elasticity from the centre adjugate and determinant
// (operators/hex8/hex8_linear_elasticity.cpp:625-760), 2x2x2 Gauss (exact):
// grad u = (sum_n u_n ghat_n) P^T with P = J^-T = adj^T/det; sigma = lambda tr(eps) I +
// 2 mu eps; y_n += w det sigma P ghat_n.  x, y interleaved xyz.
__global__ void hex8_elasticity_apply_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                             const float *__restrict__ adjg, const float *__restrict__ detg,
                                             const double mu0, const double lambda0,
                                             const float *__restrict__ coeff, const double *__restrict__ x,
                                             double *__restrict__ y) {
    const double gp[2] = {0.5 - 0.28867513459481288225, 0.5 + 0.28867513459481288225};
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        int32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = idx[k * E + e];
        double adj[9];
#pragma unroll
        for (int k = 0; k < 9; k++) adj[k] = (double)adjg[k * E + e];
        const double det = (double)detg[e];
        const double c = coeff ? (double)coeff[e] : 1.0;
        const double mu = mu0 * c, lambda = lambda0 * c;
        const double idet = 1.0 / det;
        double u[3][8], out[3][8];
#pragma unroll
        for (int n = 0; n < 8; n++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                u[d][n] = x[(ptrdiff_t)v[n] * 3 + d];
                out[d][n] = 0;
            }
        const double w = det * 0.125;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double qx = gp[q & 1], qy = gp[(q >> 1) & 1], qz = gp[q >> 2];
            // reference gradient of each component: Gh[d][a]
            double Gh[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int n = 0; n < 8; n++) {
                double g[3];
                q1_ref_grad(n, qx, qy, qz, g);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    Gh[d][0] += u[d][n] * g[0];
                    Gh[d][1] += u[d][n] * g[1];
                    Gh[d][2] += u[d][n] * g[2];
                }
            }
            // physical gradient: F[d][k] = sum_a Gh[d][a] adj[a*3+k] / det
            double F[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int k = 0; k < 3; k++)
                    F[d][k] = (Gh[d][0] * adj[0 * 3 + k] + Gh[d][1] * adj[1 * 3 + k] + Gh[d][2] * adj[2 * 3 + k]) * idet;
            const double tr = F[0][0] + F[1][1] + F[2][2];
            double Sg[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int k = 0; k < 3; k++) Sg[d][k] = mu * (F[d][k] + F[k][d]) + (d == k ? lambda * tr : 0.0);
            // back to the reference frame: R[d][a] = w * sum_k Sg[d][k] adj[a*3+k] / det
            double R[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int a = 0; a < 3; a++)
                    R[d][a] = w * idet * (Sg[d][0] * adj[a * 3 + 0] + Sg[d][1] * adj[a * 3 + 1] + Sg[d][2] * adj[a * 3 + 2]);
#pragma unroll
            for (int n = 0; n < 8; n++) {
                double g[3];
                q1_ref_grad(n, qx, qy, qz, g);
#pragma unroll
                for (int d = 0; d < 3; d++) out[d][n] += R[d][0] * g[0] + R[d][1] * g[1] + R[d][2] * g[2];
            }
        }
#pragma unroll
        for (int n = 0; n < 8; n++)
#pragma unroll
            for (int d = 0; d < 3; d++) atomicAdd(&y[(ptrdiff_t)v[n] * 3 + d], out[d][n]);
    }
}

// The same operator with eight lanes per element, one per Gauss point (small meshes, e.g. the hex8
// levels of a multigrid hierarchy, where one lane per element leaves most of the GPU idle): each
// lane evaluates its point's contribution to all 24 element rows, a butterfly over the eight lanes
// sums them, and lane n adds the rows of corner n.
__global__ void __launch_bounds__(256)
    hex8_elasticity_apply_qp_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                    const float *__restrict__ adjg, const float *__restrict__ detg,
                                    const double mu0, const double lambda0, const float *__restrict__ coeff,
                                    const double *__restrict__ x, double *__restrict__ y) {
    const double gp[2] = {0.5 - 0.28867513459481288225, 0.5 + 0.28867513459481288225};
    const int q = threadIdx.x & 7;
    // every lane of a group of eight takes part in the butterfly, also past the last element
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t - q < E * 8;
         t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t >> 3;
        const bool valid = e < E;
        const ptrdiff_t ee = valid ? e : 0;
        int32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = idx[k * E + ee];
        double adj[9];
#pragma unroll
        for (int k = 0; k < 9; k++) adj[k] = (double)adjg[k * E + ee];
        const double det = (double)detg[ee];
        const double c = coeff ? (double)coeff[ee] : 1.0;
        const double mu = mu0 * c, lambda = lambda0 * c;
        const double idet = 1.0 / det;
        const double w = det * 0.125;
        const double qx = gp[q & 1], qy = gp[(q >> 1) & 1], qz = gp[q >> 2];
        double Gh[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double g[3];
            q1_ref_grad(n, qx, qy, qz, g);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double u = x[(ptrdiff_t)v[n] * 3 + d];
                Gh[d][0] += u * g[0];
                Gh[d][1] += u * g[1];
                Gh[d][2] += u * g[2];
            }
        }
        double F[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                F[d][k] = (Gh[d][0] * adj[0 * 3 + k] + Gh[d][1] * adj[1 * 3 + k] + Gh[d][2] * adj[2 * 3 + k]) * idet;
        const double tr = F[0][0] + F[1][1] + F[2][2];
        double Sg[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int k = 0; k < 3; k++) Sg[d][k] = mu * (F[d][k] + F[k][d]) + (d == k ? lambda * tr : 0.0);
        double R[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int a = 0; a < 3; a++)
                R[d][a] = w * idet * (Sg[d][0] * adj[a * 3 + 0] + Sg[d][1] * adj[a * 3 + 1] + Sg[d][2] * adj[a * 3 + 2]);
        double out[3][8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double g[3];
            q1_ref_grad(n, qx, qy, qz, g);
#pragma unroll
            for (int d = 0; d < 3; d++) out[d][n] = R[d][0] * g[0] + R[d][1] * g[1] + R[d][2] * g[2];
        }
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
#pragma unroll
            for (int n = 0; n < 8; n++)
#pragma unroll
                for (int d = 0; d < 3; d++) out[d][n] += __shfl_xor(out[d][n], m, 8);
        if (valid) {
#pragma unroll
            for (int n = 0; n < 8; n++)
                if (n == q) {
#pragma unroll
                    for (int d = 0; d < 3; d++) atomicAdd(&y[(ptrdiff_t)v[n] * 3 + d], out[d][n]);
                }
        }
    }
}

__global__ void hex8_elasticity_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                            const float *__restrict__ adjg, const float *__restrict__ detg,
                                            const double mu0, const double lambda0, const float *__restrict__ coeff,
                                            const int block, double *__restrict__ d) {
  for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < E * 8;
       t += (ptrdiff_t)gridDim.x * blockDim.x) {
    const ptrdiff_t e = t / 8;
    const int n = (int)(t - e * 8);
    double adj[9];
    for (int k = 0; k < 9; k++) adj[k] = (double)adjg[k * E + e];
    const double det = (double)detg[e];
    const double c = coeff ? (double)coeff[e] : 1.0;
    const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
    const ptrdiff_t g = idx[n * E + e];
    if (block) {
        for (int k = 0; k < 6; k++)
            atomicAdd(&d[g * 6 + k], hex8_elasticity_entry(adj, det, mu0 * c, lambda0 * c, ca[k] * 8 + n, cb[k] * 8 + n));
    } else {
        for (int k = 0; k < 3; k++)
            atomicAdd(&d[g * 3 + k], hex8_elasticity_entry(adj, det, mu0 * c, lambda0 * c, k * 8 + n, k * 8 + n));
    }
  }
}

// ------------------------------------------------------------------ launchers
int tet4_laplacian_apply(ptrdiff_t E, const int32_t *idx, const float *fff, const double *x, double *y,
                         hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    tet4_laplacian_apply_kernel<<<grid_for(E), kBlock, 0, s>>>(E, idx, fff, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
int tet4_laplacian_diag(ptrdiff_t E, const int32_t *idx, const float *fff, double *d, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    tet4_laplacian_diag_kernel<<<grid_for(E), kBlock, 0, s>>>(E, idx, fff, d);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
int hex8_laplacian_apply(ptrdiff_t E, const int32_t *idx, const float *fff, const float *coeff, const double *x,
                         double *y, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    hex8_laplacian_apply_kernel<<<grid_for(E), kBlock, 0, s>>>(E, idx, fff, coeff, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
int hex8_laplacian_diag(ptrdiff_t E, const int32_t *idx, const float *fff, const float *coeff, double *d,
                        hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    hex8_laplacian_diag_kernel<<<grid_for(E), kBlock, 0, s>>>(E, idx, fff, coeff, d);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
int hex8_elasticity_apply(ptrdiff_t E, const int32_t *idx, const float *adj, const float *det, double mu,
                          double lambda, const float *coeff, const double *x, double *y, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    // one lane per element fills the GPU only from ~2^18 elements on; below, a lane per Gauss point
    if (E < (ptrdiff_t)1 << 18)
        hex8_elasticity_apply_qp_kernel<<<grid_for(E * 8), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, x, y);
    else
        hex8_elasticity_apply_kernel<<<grid_for(E), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
int hex8_elasticity_diag(ptrdiff_t E, const int32_t *idx, const float *adj, const float *det, double mu,
                         double lambda, const float *coeff, int block, double *d, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    hex8_elasticity_diag_kernel<<<grid_for(E * 8), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, block, d);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

}  // namespace sfem_amd
