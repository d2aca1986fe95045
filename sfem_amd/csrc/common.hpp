// Shared helpers for the gfx950 kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "sfem_amd.h"

namespace sfem_amd {

// ----------------------------------------------------------------- errors
void set_last_error(const std::string &msg);

#define SFEM_AMD_HIP_CHECK(call)                                                               \
    do {                                                                                       \
        hipError_t err__ = (call);                                                             \
        if (err__ != hipSuccess) {                                                             \
            ::sfem_amd::set_last_error(std::string(#call " failed: ") + hipGetErrorString(err__) + \
                                       " at " __FILE__ ":" + std::to_string(__LINE__));        \
            return SFEM_FAILURE;                                                               \
        }                                                                                      \
    } while (0)

#define SFEM_AMD_REQUIRE(cond, msg)                                \
    do {                                                           \
        if (!(cond)) {                                             \
            ::sfem_amd::set_last_error(std::string("sfem_amd: ") + (msg)); \
            return SFEM_FAILURE;                                   \
        }                                                          \
    } while (0)

// The library's stream for calls that pass no stream (NULL): one blocking stream per device with
// the default priority, created on first use.  Being a blocking stream it stays ordered with work on
// the legacy default stream, and unlike that stream it can be captured into a HIP graph.
hipStream_t library_stream();
inline hipStream_t as_stream(void *s) { return s ? reinterpret_cast<hipStream_t>(s) : library_stream(); }

// ----------------------------------------------------------------- structured helpers
// operators/sshex8/sshex8.hpp:18-24: local id = z*(L+1)^2 + y*(L+1) + x
__host__ __device__ constexpr inline int lidx(int L, int x, int y, int z) { return (z * (L + 1) + y) * (L + 1) + x; }
__host__ __device__ constexpr inline int nxe_of(int L) { return (L + 1) * (L + 1) * (L + 1); }
// Skeleton nodes of a macro element (on its faces, edges or corners: x, y or z in {0, L}),
// numbered in lidx order.  A line (y, z) is a skeleton line when y or z is in {0, L}: all
// of its L+1 nodes are skeleton nodes; any other line holds two (x = 0 and x = L).
__host__ __device__ constexpr inline int sk_count(int L) { return 2 * (L + 1) * (L + 1) + 4 * L * (L - 1); }
__host__ __device__ constexpr inline bool sk_line(int L, int y, int z) { return y == 0 || y == L || z == 0 || z == L; }
// slot of node (0, y, z)
__host__ __device__ constexpr inline int sk_line_base(int L, int y, int z) {
    const int N1 = L + 1;
    const int plane = z == 0 ? 0 : N1 * N1 + (z - 1) * 4 * L;
    const int in_plane = (z == 0 || z == L) ? y * N1 : (y == 0 ? 0 : N1 + (y - 1) * 2);
    return plane + in_plane;
}
// slot of skeleton node (x, y, z); -1 for interior nodes
__host__ __device__ constexpr inline int sk_slot(int L, int x, int y, int z) {
    return sk_line(L, y, z) ? sk_line_base(L, y, z) + x
                            : (x == 0 ? sk_line_base(L, y, z) : (x == L ? sk_line_base(L, y, z) + 1 : -1));
}

// hex8 corner order (python/sfem/mesh/box_mesh.py:520-534, sshex8.hpp:53-62):
// bit pattern of corner v: x = (v==1||v==2||v==5||v==6), y = (v==2||v==3||v==6||v==7), z = v>=4
__host__ __device__ constexpr inline int hex8_cx(int v) { return (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0; }
__host__ __device__ constexpr inline int hex8_cy(int v) { return (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0; }
__host__ __device__ constexpr inline int hex8_cz(int v) { return v >= 4 ? 1 : 0; }
// inverse: (x,y,z) bits -> hex8 corner id
__host__ __device__ constexpr inline int hex8_corner(int x, int y, int z) {
    return z * 4 + (y ? (x ? 2 : 3) : (x ? 1 : 0));
}

// HEX8 sides, 4 corners each in cyclic order, in the side order of python/codegen/hex8.py:44-52
// (y=0, x=1, y=1, x=0, z=0, z=1): the local side table the reference's sshex8 numbering walks
// (mesh/sfem_hex8_mesh_graph.cpp:652, filled by the absent smesh fill_local_side_table).
__host__ __device__ constexpr inline int hex8_side(int f, int i) {
    constexpr int t[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {3, 2, 1, 0}, {4, 5, 6, 7}};
    return t[f][i];
}
// side of a face-interior local node (exactly one coordinate in {0, L})
__host__ __device__ constexpr inline int hex8_side_of(int L, int x, int y, int z) {
    return y == 0 ? 0 : x == L ? 1 : y == L ? 2 : x == 0 ? 3 : z == 0 ? 4 : 5;
}

// tet4 FFF as the reference evaluates it (tet4_fff, operators/tet4/tet4_inline_cpu.hpp:224-278):
// every temporary is float32, except the volume det/6 (double constants, rounded once) and
// 1/det^2 (a double quotient, rounded once); contraction into FMA is off so that the float32
// words equal the reference's.  e1, e2, e3: edge vectors from vertex 0; aij: cofactors of J.
__host__ __device__ inline void tet4_fff_f32(const float *px, const float *py, const float *pz, float *fff) {
#pragma clang fp contract(off)
    const float e1x = px[1] - px[0], e2x = px[2] - px[0], e3x = px[3] - px[0];
    const float e1y = py[1] - py[0], e2y = py[2] - py[0], e3y = py[3] - py[0];
    const float e1z = pz[1] - pz[0], e2z = pz[2] - pz[0], e3z = pz[3] - pz[0];
    const float t_yz = e2y * e3z, p0 = e1x * t_yz;
    const float t_zy = e3y * e2z, p1 = e1x * t_zy;
    const float t_xz = e2x * e3z, p2 = t_xz * e1y;
    const float t_xy = e2x * e3y, p3 = e1z * t_xy;
    const float p4 = e3x * e2z * e1y;
    const float t_yx = e2y * e3x, p5 = e1z * t_yx;
    const double sixth = 1.0 / 6.0;
    const float vol = (float)(-sixth * p2 + sixth * p3 + sixth * p4 - sixth * p5 + sixth * p0 - sixth * p1);
    const float det = -p2 + p3 + p4 - p5 + p0 - p1;
    const float inv_det2 = (float)(1. / (det * det));
    const float a00 = t_xy - t_yx, a01 = e3x * e2z - t_xz, a02 = t_yz - t_zy;
    const float a10 = e3x * e1y - e1x * e3y, a11 = e1x * e3z - e1z * e3x, a12 = e1z * e3y - e3z * e1y;
    const float a20 = e1x * e2y - e2x * e1y, a21 = e2x * e1z - e1x * e2z, a22 = e2z * e1y - e2y * e1z;
    const float a00s = a00 * inv_det2, a01s = inv_det2 * a01, a02s = inv_det2 * a02;
    fff[0] = vol * (a00 * a00 * inv_det2 + inv_det2 * (a01 * a01) + inv_det2 * (a02 * a02));
    fff[1] = vol * (a10 * a00s + a11 * a01s + a12 * a02s);
    fff[2] = vol * (a00s * a20 + a01s * a21 + a02s * a22);
    fff[3] = vol * (inv_det2 * (a10 * a10) + inv_det2 * (a11 * a11) + inv_det2 * (a12 * a12));
    fff[4] = vol * (inv_det2 * a10 * a20 + inv_det2 * a11 * a21 + inv_det2 * a12 * a22);
    fff[5] = vol * (inv_det2 * (a20 * a20) + inv_det2 * (a21 * a21) + inv_det2 * (a22 * a22));
}

// Levels of the hierarchical sshex8 numbering, coarsest first (sshex8_hierarchical_mesh_levels,
// mesh/sfem_hex8_mesh_graph.cpp:964-992): 1, then the chain ..., L/4, L/2, L while halving stays
// even (8 -> 1 2 4 8; 12 -> 1 6 12; 6 -> 1 6).  For odd L > 1 the reference lists [1] only and its
// renumbering aborts; this library uses [1, L].  Returns the count (levels may be null).
__host__ __device__ inline int hierarchical_levels(int L, int *levels) {
    int n = 1;
    for (int l = L; l > 1 && l % 2 == 0; l /= 2) n++;
    if (n == 1 && L > 1) n = 2;
    if (levels) {
        levels[0] = 1;
        int k = n - 1;
        levels[k] = L;
        for (int l = L; l > 1 && l % 2 == 0 && k > 1; l /= 2) levels[--k] = l / 2;
    }
    return n;
}
// index k >= 1 of the coarsest level whose grid holds local node (x, y, z); 0 for macro corners
__host__ __device__ inline int hierarchical_level_index(int L, const int *levels, int nlevels, int x, int y, int z) {
    if (x % L == 0 && y % L == 0 && z % L == 0) return 0;
    for (int k = 1; k < nlevels; k++) {
        const int s = L / levels[k];
        if (x % s == 0 && y % s == 0 && z % s == 0) return k;
    }
    return nlevels - 1;
}

// 1-D Q1 integrals on [0,1]: mass, stiffness, and C[i][j] = int psi_i' psi_j.
__host__ __device__ inline double q1_mass(int i, int j) { return i == j ? 1.0 / 3.0 : 1.0 / 6.0; }
__host__ __device__ inline double q1_stiff(int i, int j) { return i == j ? 1.0 : -1.0; }
__host__ __device__ inline double q1_mixed(int i, int /*j*/) { return i ? 0.5 : -0.5; }

// I_ab(i,j) = int_{[0,1]^3} d_a phi_i d_b phi_j for hex8 corners i, j
__host__ __device__ inline double q1_grad_integral(int a, int b, int i, int j) {
    const int bi[3] = {hex8_cx(i), hex8_cy(i), hex8_cz(i)};
    const int bj[3] = {hex8_cx(j), hex8_cy(j), hex8_cz(j)};
    double r = 1.0;
    for (int d = 0; d < 3; d++) {
        double t;
        if (d == a && d == b)
            t = q1_stiff(bi[d], bj[d]);
        else if (d == a)
            t = q1_mixed(bi[d], bj[d]);
        else if (d == b)
            t = q1_mixed(bj[d], bi[d]);
        else
            t = q1_mass(bi[d], bj[d]);
        r *= t;
    }
    return r;
}

// Q1 Laplacian element matrix entry with constant metric G (packed FFF, 6 entries):
// the exact integral evaluated by operators/hex8/hex8_laplacian_inline_cpu.hpp:213-330.
__host__ __device__ inline double hex8_laplacian_entry(const double *f, int i, int j) {
    const double G[3][3] = {{f[0], f[1], f[2]}, {f[1], f[3], f[4]}, {f[2], f[4], f[5]}};
    double s = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) s += G[a][b] * q1_grad_integral(a, b, i, j);
    return s;
}

// Q1 elasticity element matrix entry, index = component*8 + node, constant Jacobian:
// K = det * [lambda M_{ci,cj} + mu M_{cj,ci} + delta_{ci,cj} mu sum_k M_kk],
// M_kl(i,j) = sum_ab P_ka P_lb I_ab(i,j), P = J^-T = adj^T / det
// (operators/hex8/hex8_linear_elasticity_inline_cpu.hpp:207-).
__host__ __device__ inline double hex8_elasticity_entry(const double *adj, double det, double mu, double lambda,
                                                        int row, int col) {
    const int ci = row / 8, i = row % 8, cj = col / 8, j = col % 8;
    double I[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) I[a][b] = q1_grad_integral(a, b, i, j);
    // P_ka = adj[a*3+k] / det
    auto M = [&](int k, int l) {
        double s = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s += adj[a * 3 + k] * adj[b * 3 + l] * I[a][b];
        return s / (det * det);
    };
    double v = lambda * M(ci, cj) + mu * M(cj, ci);
    if (ci == cj) v += mu * (M(0, 0) + M(1, 1) + M(2, 2));
    return v * det;
}

// ----------------------------------------------------------------- Chebyshev3 step
// One Chebyshev3 step on entry i given t = (A x)_i (algebra/sfem_Chebyshev3.hpp:37-52, 208-217):
// first step p = t - rhs, later ones p = beta p + (t - rhs); with a Jacobi scaling the
// (t - rhs) term is multiplied by dinv; then x -= alpha p.  Shared by the vector kernels and the
// apply epilogues so that fused and separate steps give the same words.
template <typename T>
struct ChebStep {
    const T *rhs{nullptr};
    const T *dinv{nullptr};  // optional D^-1
    T *p{nullptr};           // null: no step (a plain apply)
    T *x{nullptr};           // the operator input, updated in place
    T alpha{0}, beta{0};
    int first{0};  // p holds nothing yet (beta unused)
};
template <typename T>
__host__ __device__ __forceinline__ T cheb3_direction(T p, T rhs, T t, T beta, const T *dinv, ptrdiff_t i,
                                                      int first) {
    T r = t - rhs;
    if (dinv) r = dinv[i] * r;
    return first ? r : __builtin_fma(beta, p, r);
}

// ----------------------------------------------------------------- geometry words
// Per-element geometry is stored as float32 words (the reference's jacobian_t) or, for the FP32
// operators, as bf16 words (the upper half of the float32 word, round-to-nearest-even on the host).
__device__ __forceinline__ float bf16_to_float(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

template <typename G>
__device__ __forceinline__ double load_geom(const G *g, ptrdiff_t i);
template <>
__device__ __forceinline__ double load_geom<float>(const float *g, ptrdiff_t i) {
    return (double)g[i];
}
template <>
__device__ __forceinline__ double load_geom<uint16_t>(const uint16_t *g, ptrdiff_t i) {
    return (double)bf16_to_float(g[i]);
}

// ----------------------------------------------------------------- workgroup barrier for LDS
// Orders LDS accesses only: __syncthreads() is a workgroup fence over all address spaces, which on
// this chip waits for every outstanding global load and store (vmcnt(0)) before the barrier.
// Kernels whose barriers only separate LDS writes from LDS reads use this one, so global stores
// issued before it stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD b % 8), each with its
// own L2.  This renumbering gives XCD k the k-th contiguous range of the grid, so workgroups on
// neighbouring elements (consecutive ids, sharing boundary nodes and the cache lines of their
// vector entries) meet in one L2 instead of writing back partial lines from several.
__device__ __forceinline__ ptrdiff_t xcd_contiguous_block(ptrdiff_t b, ptrdiff_t n) {
    const ptrdiff_t per = n / 8, rem = n % 8, xcd = b % 8;
    return xcd * per + (xcd < rem ? xcd : rem) + b / 8;
}

// ----------------------------------------------------------------- launch helpers
constexpr int kBlock = 256;
inline unsigned grid_for(ptrdiff_t n, int block = kBlock, unsigned cap = 1u << 20) {
    ptrdiff_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (ptrdiff_t)cap) g = cap;
    return (unsigned)g;
}

}  // namespace sfem_amd
