// gfx950 kernels for tet4 and hex8 meshes: one element per lane.
//
// Device layout: idx int32 SoA [nv][E] and geometry SoA [k][E] (the layout of the
// reference's device path, operators/cuda/cu_laplacian.hpp fff_stride), so a wave
// reads 64 consecutive words of every row: fully coalesced.  Scatter with FP64
// atomics, as cu_tet4_laplacian_apply_kernel (operators/tet4/cuda/cu_tet4_laplacian.cu:11-44).
#include <type_traits>

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
// (operators/hex8/hex8_linear_elasticity.cpp:625-760), 2x2x2 Gauss (exact for a constant Jacobian):
// grad u = (sum_n u_n ghat_n) P^T with P = J^-T = adj^T/det; sigma = lambda tr(eps) I + 2 mu eps;
// y_n += w det sigma P ghat_n.  x, y interleaved xyz.
//
// Sum-factorised, one element per lane.  With the corners as a 2x2x2 tensor u[iz][iy][ix] and the
// 1-D Gauss points g_0, g_1, the reference derivative along x at point (qz, qy, qx) is the bilinear
// interpolation in (z, y) of the four x-differences u[iz][iy][1] - u[iz][iy][0], the same for both
// qx: each of the 9 (component, direction) gradients has 4 distinct values over the 8 points
// (16 operations each instead of 64).  The transpose runs the same steps backwards: the point
// values of each (component, direction) are first summed over the direction's own index, then
// spread bilinearly and differenced onto the corners.  About 1,000 operations per element against
// 3,700 for the plain point loop; T is the value type (double, or float for the FP32 operator with
// bf16 geometry words G).
template <typename T>
__device__ __forceinline__ void sf_interp2(const T a00, const T a01, const T a10, const T a11, const T g0, const T g1,
                                           T out[2][2]) {
    // out[q1][q2] = sum_ij B[q1][i] B[q2][j] a_ij with B[q][0] = 1 - g_q, B[q][1] = g_q
    const T r00 = a00 + g0 * (a01 - a00), r01 = a00 + g1 * (a01 - a00);
    const T r10 = a10 + g0 * (a11 - a10), r11 = a10 + g1 * (a11 - a10);
    out[0][0] = r00 + g0 * (r10 - r00);
    out[1][0] = r00 + g1 * (r10 - r00);
    out[0][1] = r01 + g0 * (r11 - r01);
    out[1][1] = r01 + g1 * (r11 - r01);
}
template <typename T>
__device__ __forceinline__ void sf_interp2_t(const T r[2][2], const T g0, const T g1, T &a00, T &a01, T &a10,
                                             T &a11) {
    // transpose of sf_interp2: a_ij = sum_{q1 q2} B[q1][i] B[q2][j] r[q1][q2]; first over q2:
    // H[q1][1] = sum_q2 g_q2 r[q1][q2], H[q1][0] = sum_q2 r[q1][q2] - H[q1][1]
    const T h01 = g0 * r[0][0] + g1 * r[0][1], h00 = (r[0][0] + r[0][1]) - h01;
    const T h11 = g0 * r[1][0] + g1 * r[1][1], h10 = (r[1][0] + r[1][1]) - h11;
    // then over q1 the same way
    a10 = g0 * h00 + g1 * h10;
    a11 = g0 * h01 + g1 * h11;
    a00 = (h00 + h10) - a10;
    a01 = (h01 + h11) - a11;
}

// One element: corner values load(n, d) (hex8 corner n, component d) in, store(n, d, v) of the 24
// products out; a64/det: the element's adjugate and determinant, c its coefficient.
template <typename T, typename Load, typename Store>
__device__ __forceinline__ void hex8_elasticity_sf_element(const double a64[9], const double det, const double c,
                                                           const double mu0, const double lambda0, Load load,
                                                           Store store) {
    // hex8 corner n -> tensor position iz*4 + iy*2 + ix
    constexpr int kT[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const T g0 = (T)(0.5 - 0.28867513459481288225), g1 = (T)(0.5 + 0.28867513459481288225);
    // P[a][k] = adj[a][k] / det for the physical gradient; the weight det/8 and the 1/det of the way
    // back leave adj/8 on the stress side, folded into mu and lambda
    const double idet = 1.0 / det;
    T P[9], A[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        P[k] = (T)(a64[k] * idet);
        A[k] = (T)a64[k];
    }
    const T mu = (T)(mu0 * c * 0.125), lambda = (T)(lambda0 * c * 0.125);
    // gx[c][qz][qy], gy[c][qz][qx], gz[c][qy][qx]
    T gx[3][2][2], gy[3][2][2], gz[3][2][2];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T u[8];
#pragma unroll
        for (int n = 0; n < 8; n++) u[kT[n]] = load(n, d);
        // x-differences at (iz, iy), interpolated in (z, y)
        sf_interp2<T>(u[1] - u[0], u[3] - u[2], u[5] - u[4], u[7] - u[6], g0, g1, gx[d]);
        // y-differences at (iz, ix), interpolated in (z, x)
        sf_interp2<T>(u[2] - u[0], u[3] - u[1], u[6] - u[4], u[7] - u[5], g0, g1, gy[d]);
        // z-differences at (iy, ix), interpolated in (y, x)
        sf_interp2<T>(u[4] - u[0], u[5] - u[1], u[6] - u[2], u[7] - u[3], g0, g1, gz[d]);
    }
    // point values of R[d][a] = sum_k S[d][k] adj[a][k], summed over the direction's own index
    T rx[3][2][2], ry[3][2][2], rz[3][2][2];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) rx[d][i][j] = ry[d][i][j] = rz[d][i][j] = 0;
#pragma unroll
    for (int qz = 0; qz < 2; qz++)
#pragma unroll
        for (int qy = 0; qy < 2; qy++)
#pragma unroll
            for (int qx = 0; qx < 2; qx++) {
                T F[3][3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const T h0 = gx[d][qz][qy], h1 = gy[d][qz][qx], h2 = gz[d][qy][qx];
#pragma unroll
                    for (int k = 0; k < 3; k++) F[d][k] = h0 * P[0 * 3 + k] + h1 * P[1 * 3 + k] + h2 * P[2 * 3 + k];
                }
                const T ltr = lambda * (F[0][0] + F[1][1] + F[2][2]);
                T S[3][3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    S[d][d] = 2 * mu * F[d][d] + ltr;
#pragma unroll
                    for (int k = d + 1; k < 3; k++) S[d][k] = S[k][d] = mu * (F[d][k] + F[k][d]);
                }
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    rx[d][qz][qy] += S[d][0] * A[0] + S[d][1] * A[1] + S[d][2] * A[2];
                    ry[d][qz][qx] += S[d][0] * A[3] + S[d][1] * A[4] + S[d][2] * A[5];
                    rz[d][qy][qx] += S[d][0] * A[6] + S[d][1] * A[7] + S[d][2] * A[8];
                }
            }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T o[8];
        T a00, a01, a10, a11;
        sf_interp2_t<T>(rx[d], g0, g1, a00, a01, a10, a11);  // at (iz, iy): +x corner gets +, -x gets -
        o[1] = a00, o[0] = -a00, o[3] = a01, o[2] = -a01, o[5] = a10, o[4] = -a10, o[7] = a11, o[6] = -a11;
        sf_interp2_t<T>(ry[d], g0, g1, a00, a01, a10, a11);  // at (iz, ix)
        o[2] += a00, o[0] -= a00, o[3] += a01, o[1] -= a01, o[6] += a10, o[4] -= a10, o[7] += a11, o[5] -= a11;
        sf_interp2_t<T>(rz[d], g0, g1, a00, a01, a10, a11);  // at (iy, ix)
        o[4] += a00, o[0] -= a00, o[5] += a01, o[1] -= a01, o[6] += a10, o[2] -= a10, o[7] += a11, o[3] -= a11;
#pragma unroll
        for (int n = 0; n < 8; n++) store(n, d, o[kT[n]]);
    }
}

// Element-wise: one element per lane, products added to y with atomics.
template <typename T, typename G>
__global__ void __launch_bounds__(256)
    hex8_elasticity_apply_sf_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx, const G *__restrict__ adjg,
                                    const G *__restrict__ detg, const double mu0, const double lambda0,
                                    const float *__restrict__ coeff, const T *__restrict__ x, T *__restrict__ y) {
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (ptrdiff_t)gridDim.x * blockDim.x) {
        int32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = idx[k * E + e];
        double a64[9];
#pragma unroll
        for (int k = 0; k < 9; k++) a64[k] = load_geom<G>(adjg, k * E + e);
        hex8_elasticity_sf_element<T>(
            a64, load_geom<G>(detg, e), coeff ? (double)coeff[e] : 1.0, mu0, lambda0,
            [&](int n, int d) { return x[(ptrdiff_t)v[n] * 3 + d]; },
            [&](int n, int d, T val) { atomicAdd(&y[(ptrdiff_t)v[n] * 3 + d], val); });
    }
}

// Patch-wise (ElementPatches): one workgroup per patch of up to kPatch elements.  The patch's nodes
// are gathered into LDS once (x), the elements accumulate into LDS (y, LDS atomics), then each patch
// node is either added to y (the patch's own nodes: no other workgroup touches them) or written to
// its node-sorted partial slot (nodes shared with other patches; summed by skeleton_gather).  The
// geometry and coefficient words are stored in patch-slot order, so every read of them coalesces.
// The LDS accumulator is FP64 for both value types: ds_add_f64 keeps pace with the arithmetic, while
// ds_add_f32 measured 7x the LDS busy cycles and 100x the LDS wait of the whole FP64 kernel here.
template <typename T, typename G, int kPatch>
__global__ void __launch_bounds__(256)
    hex8_elasticity_patch_kernel(const int32_t *__restrict__ count, const uint16_t *__restrict__ lconn,
                                 const int32_t *__restrict__ node_off, const int32_t *__restrict__ nodes,
                                 const int32_t *__restrict__ bslot, const ptrdiff_t Epad, const G *__restrict__ adjp,
                                 const G *__restrict__ detp, const float *__restrict__ coeffp, const double mu0,
                                 const double lambda0, const T *__restrict__ x, T *__restrict__ y,
                                 T *__restrict__ part) {
    extern __shared__ __align__(16) unsigned char patch_lds[];
    // workgroups go to the 8 XCDs round-robin: give XCD k the k-th contiguous range of patches, so
    // neighbouring patches (consecutive in growth order, sharing boundary nodes) meet in one L2
    const ptrdiff_t b = blockIdx.x, per = gridDim.x / 8, rem = gridDim.x % 8;
    const int xcd = (int)(b % 8);
    const ptrdiff_t p = xcd * per + (xcd < rem ? xcd : rem) + b / 8;
    const int off = node_off[p], nloc = node_off[p + 1] - off;
    double *sy = reinterpret_cast<double *>(patch_lds);
    T *sx = reinterpret_cast<T *>(sy + 3 * nloc);
    // NB patch nodes per thread at a time: their ids, then their x values, each batch issued together
    constexpr int NB = 4;
    for (int i0 = threadIdx.x; i0 < nloc; i0 += NB * blockDim.x) {
        ptrdiff_t gg[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int i = i0 + b * blockDim.x;
            gg[b] = nodes[off + (i < nloc ? i : 0)];
        }
        T xv[NB][3];
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) xv[b][c] = x[gg[b] * 3 + c];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int i = i0 + b * blockDim.x;
            if (i < nloc)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    sx[i * 3 + c] = xv[b][c];
                    sy[i * 3 + c] = 0;
                }
        }
    }
    __syncthreads();
    const int cnt = count[p];
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        const ptrdiff_t slot = p * kPatch + t;
        int l[8];
#pragma unroll
        for (int k = 0; k < 8; k++) l[k] = lconn[(p * 8 + k) * kPatch + t];
        double a64[9];
#pragma unroll
        for (int k = 0; k < 9; k++) a64[k] = load_geom<G>(adjp, k * Epad + slot);
        hex8_elasticity_sf_element<T>(
            a64, load_geom<G>(detp, slot), coeffp ? (double)coeffp[slot] : 1.0, mu0, lambda0,
            [&](int n, int d) { return sx[l[n] * 3 + d]; },
            [&](int n, int d, T val) { atomicAdd(&sy[l[n] * 3 + d], (double)val); });
    }
    __syncthreads();
    // the same batches: ids and boundary slots, then the old y (read for every node of the batch: a
    // valid entry, used for the patch's own nodes), then the stores
    for (int i0 = threadIdx.x; i0 < nloc; i0 += NB * blockDim.x) {
        ptrdiff_t gg[NB], bb[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int i = i0 + b * blockDim.x, ii = i < nloc ? i : 0;
            gg[b] = nodes[off + ii];
            bb[b] = bslot[off + ii];
        }
        T yo[NB][3];
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) yo[b][c] = y[gg[b] * 3 + c];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int i = i0 + b * blockDim.x;
            if (i >= nloc) continue;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T v = (T)sy[i * 3 + c];
                if (bb[b] < 0)
                    y[gg[b] * 3 + c] = yo[b][c] + v;
                else
                    part[bb[b] * 3 + c] = v;
            }
        }
    }
}

// out[k][s] = in[k][perm[s]] (0 where perm[s] < 0): per-element rows in patch-slot order
template <typename G>
__global__ void patch_gather_rows_kernel(const int rows, const ptrdiff_t E, const ptrdiff_t Epad,
                                         const int32_t *__restrict__ perm, const G *__restrict__ in,
                                         G *__restrict__ out) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows * Epad;
         i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t k = i / Epad, s = i - k * Epad;
        const int32_t e = perm[s];
        out[i] = e >= 0 ? in[k * E + e] : G(0);
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

template <typename T, typename G>
__global__ void hex8_elasticity_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                            const G *__restrict__ adjg, const G *__restrict__ detg,
                                            const double mu0, const double lambda0, const float *__restrict__ coeff,
                                            const int block, T *__restrict__ d) {
  for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < E * 8;
       t += (ptrdiff_t)gridDim.x * blockDim.x) {
    const ptrdiff_t e = t / 8;
    const int n = (int)(t - e * 8);
    double adj[9];
    for (int k = 0; k < 9; k++) adj[k] = load_geom<G>(adjg, k * E + e);
    const double det = load_geom<G>(detg, e);
    const double c = coeff ? (double)coeff[e] : 1.0;
    const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
    const ptrdiff_t g = idx[n * E + e];
    if (block) {
        for (int k = 0; k < 6; k++)
            atomicAdd(&d[g * 6 + k],
                      (T)hex8_elasticity_entry(adj, det, mu0 * c, lambda0 * c, ca[k] * 8 + n, cb[k] * 8 + n));
    } else {
        for (int k = 0; k < 3; k++)
            atomicAdd(&d[g * 3 + k], (T)hex8_elasticity_entry(adj, det, mu0 * c, lambda0 * c, k * 8 + n, k * 8 + n));
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
template <typename T, typename G>
int hex8_elasticity_apply(ptrdiff_t E, const int32_t *idx, const G *adj, const G *det, double mu, double lambda,
                          const float *coeff, const T *x, T *y, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    // one lane per element fills the GPU only from ~2^18 elements on; below, the FP64 operator runs
    // a lane per Gauss point
    if constexpr (std::is_same<T, double>::value && std::is_same<G, float>::value) {
        if (E < (ptrdiff_t)1 << 18) {
            hex8_elasticity_apply_qp_kernel<<<grid_for(E * 8), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, x,
                                                                               y);
            SFEM_AMD_HIP_CHECK(hipGetLastError());
            return SFEM_SUCCESS;
        }
    }
    hex8_elasticity_apply_sf_kernel<T, G><<<grid_for(E), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T, typename G>
int hex8_elasticity_diag(ptrdiff_t E, const int32_t *idx, const G *adj, const G *det, double mu, double lambda,
                         const float *coeff, int block, T *d, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    hex8_elasticity_diag_kernel<T, G>
        <<<grid_for(E * 8), kBlock, 0, s>>>(E, idx, adj, det, mu, lambda, coeff, block, d);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T, typename G>
int hex8_elasticity_apply_patches(const PatchView &pv, const G *adjp, const G *detp, const float *coeffp, double mu,
                                  double lambda, const T *x, T *y, T *part, hipStream_t s) {
    if (pv.npatch == 0) return SFEM_SUCCESS;
    const size_t lds = hex8_patch_lds_bytes(pv.maxloc, sizeof(T));
    if (lds > kPatchLdsLimit) {
        set_last_error("hex8 elasticity patches: patch node count exceeds the LDS budget");
        return SFEM_FAILURE;
    }
    hex8_elasticity_patch_kernel<T, G, kPatchElements><<<(unsigned)pv.npatch, 256, lds, s>>>(
        pv.count, pv.lconn, pv.node_off, pv.nodes, pv.bslot, pv.npatch * kPatchElements, adjp, detp, coeffp, mu,
        lambda, x, y, part);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename G>
int patch_gather_rows(int rows, ptrdiff_t E, ptrdiff_t Epad, const int32_t *perm, const G *in, G *out,
                      hipStream_t s) {
    if (rows * Epad == 0) return SFEM_SUCCESS;
    patch_gather_rows_kernel<G><<<grid_for(rows * Epad), kBlock, 0, s>>>(rows, E, Epad, perm, in, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template int hex8_elasticity_apply_patches<double, float>(const PatchView &, const float *, const float *,
                                                          const float *, double, double, const double *, double *,
                                                          double *, hipStream_t);
template int hex8_elasticity_apply_patches<float, uint16_t>(const PatchView &, const uint16_t *, const uint16_t *,
                                                            const float *, double, double, const float *, float *,
                                                            float *, hipStream_t);
template int patch_gather_rows<float>(int, ptrdiff_t, ptrdiff_t, const int32_t *, const float *, float *,
                                      hipStream_t);
template int patch_gather_rows<uint16_t>(int, ptrdiff_t, ptrdiff_t, const int32_t *, const uint16_t *, uint16_t *,
                                         hipStream_t);
// FP64 operator on float32 geometry words; FP32 operator on bf16 words
template int hex8_elasticity_apply<double, float>(ptrdiff_t, const int32_t *, const float *, const float *, double,
                                                  double, const float *, const double *, double *, hipStream_t);
template int hex8_elasticity_apply<float, uint16_t>(ptrdiff_t, const int32_t *, const uint16_t *, const uint16_t *,
                                                    double, double, const float *, const float *, float *,
                                                    hipStream_t);
template int hex8_elasticity_diag<double, float>(ptrdiff_t, const int32_t *, const float *, const float *, double,
                                                 double, const float *, int, double *, hipStream_t);
template int hex8_elasticity_diag<float, uint16_t>(ptrdiff_t, const int32_t *, const uint16_t *, const uint16_t *,
                                                   double, double, const float *, int, float *, hipStream_t);

}  // namespace sfem_amd
