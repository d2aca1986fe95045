// gfx950 kernels for semi-structured hex8 (sshex8) macro elements.
//
// Device layout (set up by mesh.cpp from the reference's SoA [nxe][E]):
//   idx   int32 [E][nxe]   element-major, so one macro element's (L+1)^3 node ids are
//                          one contiguous, coalesced block;
//   fff   float/bf16 [6][E] SoA per-macro FFF (operators/sshex8/sshex8_laplacian.cpp:698-769
//                          reads the macro FFF, scales by h = 1/L: hex8_sub_fff_0 :19-28);
//   coeff float [E]        optional per-macro multiplier.
// Vectors are T (double for the plain method, float for the FP32 preconditioner).
#include "common.hpp"
#include "sshex8_kernels.hpp"

#include <hip/hip_bf16.h>

#include <type_traits>
#include <utility>

namespace sfem_amd {

template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) {
    atomicAdd(p, v);
}

constexpr int round64(int n) { return ((n + 63) / 64) * 64; }

template <int L>
struct LapShape {
    static constexpr int N1 = L + 1;
    static constexpr int NXE = N1 * N1 * N1;
    static constexpr int LINES = N1 * N1;
    static constexpr int EPB = LINES >= 256 ? 1 : 256 / LINES;  // macro elements per block
    static constexpr int BLOCK = round64(EPB * LINES);
};

// Computes the 8x8 sub-element matrices of the block's macro elements into sK.
template <int L, int EPB, int BLOCK, typename G>
__device__ __forceinline__ void lap_block_matrices(ptrdiff_t E, ptrdiff_t e0, const G *__restrict__ fff,
                                                   ptrdiff_t fstride, const float *__restrict__ coeff,
                                                   const double *__restrict__ Kem, double (*sK)[64]) {
    const double h = 1.0 / L;
    for (int k = threadIdx.x; k < EPB * 64; k += BLOCK) {
        const int le = k >> 6, ij = k & 63;
        const ptrdiff_t e = e0 + le;
        double v = 0;
        if (e < E && Kem) {
            // stored sub-element matrices [E][64], row-major in hex8 corner order
            // (operators/stencil/sshex8_stencil_element_matrix_apply.cpp:10-78)
            v = Kem[e * 64 + ij];
        } else if (e < E) {
            double f[6];
            const double c = (coeff ? (double)coeff[e] : 1.0) * h;
#pragma unroll
            for (int q = 0; q < 6; q++) f[q] = load_geom<G>(fff, q * fstride + e) * c;
            v = hex8_laplacian_entry(f, ij >> 3, ij & 7);
        }
        sK[le][ij] = v;
    }
}

// y += A x for the sshex8 Laplacian.  One thread per x-line (y,z) of a macro element;
// the line's node values sit in LDS, a 3x3x3 register window slides along x.  The
// node's stencil is split into S_a (micro elements whose lower x-corner is the node) and
// S_b (upper x-corner), each masked by which y/z micro elements exist inside the macro
// element, so x = 0 and x = L (skeleton) need no divergent code: the whole wave is at
// the same x.  Interior nodes are owned by one macro element: their y values are loaded
// together with x (before the barrier), so the sweep only stores.
// Skeleton nodes are shared: their element contributions go to the element's slots of
// `part` (plain stores, no atomics) and skeleton_gather_kernel sums them per node in a
// fixed order, so the result is bitwise reproducible.
template <typename T, typename G, int L>
__global__ void __launch_bounds__(LapShape<L>::BLOCK)
    sshex8_laplacian_apply_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx, const G *__restrict__ fff,
                                  const ptrdiff_t fstride, const float *__restrict__ coeff,
                                  const double *__restrict__ Kem, const T *__restrict__ x, T *__restrict__ y,
                                  T *__restrict__ part, const int32_t *__restrict__ sk_pos, const int overwrite) {
    using S = LapShape<L>;
    constexpr int N1 = S::N1, NXE = S::NXE, LINES = S::LINES, EPB = S::EPB, NSK = sk_count(L);
    __shared__ T su[EPB][NXE];
    __shared__ double sK[EPB][64];

    const int tid = threadIdx.x;
    const int le = tid / LINES;
    const int line = tid - le * LINES;
    const int yy = line % N1, zz = line / N1;
    const ptrdiff_t e0 = xcd_contiguous_block(blockIdx.x, gridDim.x) * EPB;
    const ptrdiff_t e = e0 + le;
    const bool active = (le < EPB) && (e < E);

    const bool skl = sk_line(L, yy, zz);
    int32_t gid[N1];
    T old[N1];  // y of the interior nodes of an interior line, loaded early
    if (active) {
        const int32_t *row = idx + e * NXE + lidx(L, 0, yy, zz);
#pragma unroll
        for (int i = 0; i < N1; i++) gid[i] = row[i];
#pragma unroll
        for (int i = 0; i < N1; i++) su[le][lidx(L, i, yy, zz)] = x[gid[i]];
        if (!skl) {
#pragma unroll
            for (int i = 1; i < L; i++) old[i] = overwrite ? T(0) : y[gid[i]];
        }
    }
    lap_block_matrices<L, EPB, S::BLOCK, G>(E, e0, fff, fstride, coeff, Kem, sK);
    __syncthreads();
    if (!active) return;

    // masked stencil halves: [dz+1][dy+1][dx'] with dx' = dx (S_a, dx in {0,1}) or dx+1 (S_b)
    T Sa[3][3][2], Sb[3][3][2];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                Sa[a][b][c] = 0;
                Sb[a][b][c] = 0;
            }
    const double *K = sK[le];
#pragma unroll
    for (int cz = 0; cz < 2; cz++) {
        const bool okz = cz ? (zz > 0) : (zz < L);
#pragma unroll
        for (int cy = 0; cy < 2; cy++) {
            const bool oky = cy ? (yy > 0) : (yy < L);
            const double m = (okz && oky) ? 1.0 : 0.0;
#pragma unroll
            for (int dz = -cz; dz <= 1 - cz; dz++)
#pragma unroll
                for (int dy = -cy; dy <= 1 - cy; dy++) {
                    const int ra = hex8_corner(0, cy, cz), rb = hex8_corner(1, cy, cz);
#pragma unroll
                    for (int dx = 0; dx < 2; dx++) {
                        Sa[dz + 1][dy + 1][dx] += (T)(m * K[ra * 8 + hex8_corner(dx, cy + dy, cz + dz)]);
                        Sb[dz + 1][dy + 1][dx] += (T)(m * K[rb * 8 + hex8_corner(dx, cy + dy, cz + dz)]);
                    }
                }
        }
    }

    // sliding window: w[s][dz+1][dy+1], s = 0: x-1, 1: x, 2: x+1
    const int yl = yy > 0 ? yy - 1 : 0, yh = yy < L ? yy + 1 : L;
    const int zl = zz > 0 ? zz - 1 : 0, zh = zz < L ? zz + 1 : L;
    const int ys[3] = {yl, yy, yh}, zs[3] = {zl, zz, zh};
    const T *u = su[le];
    T w[3][3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            w[0][a][b] = 0;
            w[1][a][b] = u[lidx(L, 0, ys[b], zs[a])];
            w[2][a][b] = u[lidx(L, 1, ys[b], zs[a])];
        }

    // part positions of this line's skeleton nodes (all L+1 on a skeleton line, else x = 0, L);
    // part == null: skeleton contributions are added to y atomically instead (accumulating apply)
    const int32_t *pp = sk_pos + e * NSK + sk_line_base(L, yy, zz);
    int32_t ppos[N1];
    if (part) {
        if (skl) {
#pragma unroll
            for (int i = 0; i <= L; i++) ppos[i] = pp[i];
        } else {
            ppos[0] = pp[0];
            ppos[L] = pp[1];
        }
    }
#pragma unroll
    for (int i = 0; i <= L; i++) {
        T v = 0;
        if (i < L) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    v = fma(Sa[a][b][0], w[1][a][b], v);
                    v = fma(Sa[a][b][1], w[2][a][b], v);
                }
        }
        if (i > 0) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    v = fma(Sb[a][b][0], w[0][a][b], v);
                    v = fma(Sb[a][b][1], w[1][a][b], v);
                }
        }
        if (skl || i == 0 || i == L) {
            if (part)
                part[ppos[i]] = v;
            else
                atomic_add(&y[gid[i]], v);
        } else {
            y[gid[i]] = old[i] + v;
        }
        // slide
        if (i < L) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    w[0][a][b] = w[1][a][b];
                    w[1][a][b] = w[2][a][b];
                    w[2][a][b] = (i + 2 <= L) ? u[lidx(L, i + 2 <= L ? i + 2 : L, ys[b], zs[a])] : T(0);
                }
        }
    }

}

// The same operator with the block's memory traffic staged through LDS (levels whose block
// tile fits in 64 KB).  A thread per x-line makes each of the line's loads and stores a
// 64-lane access spread over ~30 cache lines (lanes one line apart); here the block's node
// slots are read and written with consecutive threads on consecutive slots of the
// element-major connectivity, whose ids run in long consecutive stretches, instead:
//   1. per slot: its id, x, and the old y (interior nodes, accumulating) or the skeleton slot
//      position, into LDS (one unrolled batch of loads per thread);
//   2. the line sweep of the kernel above, from LDS x into the LDS results (each node has one
//      line thread, so no two threads touch a result);
//   3. per slot: interior nodes store their y, skeleton nodes their partial (or, without
//      partials, add to y atomically).
template <int L>
struct LapStagedShape {
    static constexpr int N1 = L + 1, NXE = N1 * N1 * N1, LINES = N1 * N1;
    static constexpr int EPB = LapShape<L>::EPB, BLOCK = LapShape<L>::BLOCK;
    static constexpr int SLOTS = EPB * NXE, PER_THREAD = (SLOTS + BLOCK - 1) / BLOCK;
    template <typename T>
    static constexpr size_t lds_bytes() {
        return (size_t)SLOTS * (2 * sizeof(T) + sizeof(int32_t)) + (size_t)EPB * 64 * sizeof(double);
    }
};

// Used for FP32 vectors: at 113 VGPRs the staged kernel runs 4 waves per SIMD and beats the
// line-thread loads (L=8 on 32^3: 165 vs 189 us); in FP64 both kernels hold ~200 VGPRs (2 waves
// per SIMD, the window and the masked stencil halves) and the three phases of the staged one
// overlap worse (280 vs 242 us), so FP64 keeps the line-thread kernel.
template <typename T, int L>
constexpr bool lap_staged_used() {
    return std::is_same<T, float>::value && LapStagedShape<L>::template lds_bytes<T>() <= 64 * 1024;
}

template <typename T, typename G, int L>
__global__ void __launch_bounds__(LapShape<L>::BLOCK)
    sshex8_laplacian_apply_staged_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx, const G *__restrict__ fff,
                                         const ptrdiff_t fstride, const float *__restrict__ coeff,
                                         const double *__restrict__ Kem, const T *__restrict__ x, T *__restrict__ y,
                                         T *__restrict__ part, const int32_t *__restrict__ sk_pos,
                                         const int overwrite) {
    using S = LapStagedShape<L>;
    constexpr int N1 = S::N1, NXE = S::NXE, LINES = S::LINES, EPB = S::EPB, NSK = sk_count(L);
    __shared__ T su[EPB][NXE];      // x of the block's nodes
    __shared__ T sy[EPB][NXE];      // old y (interior, accumulating) or 0, then the results
    __shared__ int32_t st[EPB][NXE];  // store target: node id (interior; skeleton without partials) or slot position
    __shared__ double sK[EPB][64];

    const int tid = threadIdx.x;
    const ptrdiff_t e0 = xcd_contiguous_block(blockIdx.x, gridDim.x) * EPB;
    const ptrdiff_t nslots = (E - e0 < EPB ? E - e0 : EPB) * (ptrdiff_t)NXE;
    T *su0 = &su[0][0], *sy0 = &sy[0][0];
    int32_t *st0 = &st[0][0];

    // 1. ids, then x and old y or slot positions, of this thread's slots
    {
        int32_t id[S::PER_THREAD];
#pragma unroll
        for (int k = 0; k < S::PER_THREAD; k++) {
            const int sl = tid + k * S::BLOCK;
            id[k] = sl < nslots ? idx[e0 * NXE + sl] : 0;
        }
        T xv[S::PER_THREAD], yv[S::PER_THREAD];
        int32_t tg[S::PER_THREAD];
#pragma unroll
        for (int k = 0; k < S::PER_THREAD; k++) {
            const int sl = tid + k * S::BLOCK;
            const int le = sl / NXE, n = sl - le * NXE;
            const int cx = n % N1, cy = (n / N1) % N1, cz = n / (N1 * N1);
            const int slot = sk_slot(L, cx, cy, cz);
            const bool ok = sl < nslots;
            xv[k] = ok ? x[id[k]] : T(0);
            yv[k] = (ok && slot < 0 && !overwrite) ? y[id[k]] : T(0);
            tg[k] = (ok && slot >= 0 && part) ? sk_pos[(e0 + le) * NSK + slot] : id[k];
        }
#pragma unroll
        for (int k = 0; k < S::PER_THREAD; k++) {
            const int sl = tid + k * S::BLOCK;
            if (sl < S::SLOTS) {
                su0[sl] = xv[k];
                sy0[sl] = yv[k];
                st0[sl] = tg[k];
            }
        }
    }
    lap_block_matrices<L, EPB, S::BLOCK, G>(E, e0, fff, fstride, coeff, Kem, sK);
    __syncthreads();

    // 2. the line sweep
    const int le = tid / LINES, line = tid - le * LINES;
    const int yy = line % N1, zz = line / N1;
    if (le < EPB && e0 + le < E) {
        T Sa[3][3][2], Sb[3][3][2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    Sa[a][b][c] = 0;
                    Sb[a][b][c] = 0;
                }
        const double *K = sK[le];
#pragma unroll
        for (int cz = 0; cz < 2; cz++) {
            const bool okz = cz ? (zz > 0) : (zz < L);
#pragma unroll
            for (int cy = 0; cy < 2; cy++) {
                const bool oky = cy ? (yy > 0) : (yy < L);
                const double m = (okz && oky) ? 1.0 : 0.0;
#pragma unroll
                for (int dz = -cz; dz <= 1 - cz; dz++)
#pragma unroll
                    for (int dy = -cy; dy <= 1 - cy; dy++) {
                        const int ra = hex8_corner(0, cy, cz), rb = hex8_corner(1, cy, cz);
#pragma unroll
                        for (int dx = 0; dx < 2; dx++) {
                            Sa[dz + 1][dy + 1][dx] += (T)(m * K[ra * 8 + hex8_corner(dx, cy + dy, cz + dz)]);
                            Sb[dz + 1][dy + 1][dx] += (T)(m * K[rb * 8 + hex8_corner(dx, cy + dy, cz + dz)]);
                        }
                    }
            }
        }
        const int yl = yy > 0 ? yy - 1 : 0, yh = yy < L ? yy + 1 : L;
        const int zl = zz > 0 ? zz - 1 : 0, zh = zz < L ? zz + 1 : L;
        const int ys[3] = {yl, yy, yh}, zs[3] = {zl, zz, zh};
        const T *__restrict__ u = su[le];
        T *__restrict__ out = sy[le];
        T res[N1];
        T w[3][3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                w[0][a][b] = 0;
                w[1][a][b] = u[lidx(L, 0, ys[b], zs[a])];
                w[2][a][b] = u[lidx(L, 1, ys[b], zs[a])];
            }
#pragma unroll
        for (int i = 0; i <= L; i++) {
            T v = 0;
            if (i < L) {
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        v = fma(Sa[a][b][0], w[1][a][b], v);
                        v = fma(Sa[a][b][1], w[2][a][b], v);
                    }
            }
            if (i > 0) {
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        v = fma(Sb[a][b][0], w[0][a][b], v);
                        v = fma(Sb[a][b][1], w[1][a][b], v);
                    }
            }
            res[i] = v;
            if (i < L) {
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        w[0][a][b] = w[1][a][b];
                        w[1][a][b] = w[2][a][b];
                        w[2][a][b] = (i + 2 <= L) ? u[lidx(L, i + 2 <= L ? i + 2 : L, ys[b], zs[a])] : T(0);
                    }
            }
        }
        // the line's results after the sweep (no LDS store inside it, so the window may be re-read)
#pragma unroll
        for (int i = 0; i <= L; i++) {
            const int n = lidx(L, i, yy, zz);
            out[n] += res[i];
        }
    }
    __syncthreads();

    // 3. the stores
#pragma unroll
    for (int k = 0; k < S::PER_THREAD; k++) {
        const int sl = tid + k * S::BLOCK;
        if (sl < nslots) {
            const int n = sl % NXE;
            const int cx = n % N1, cy = (n / N1) % N1, cz = n / (N1 * N1);
            const bool skel = sk_slot(L, cx, cy, cz) >= 0;
            const T v = sy0[sl];
            const int32_t t = st0[sl];
            if (!skel)
                y[t] = v;
            else if (part)
                part[t] = v;
            else
                atomic_add(&y[t], v);
        }
    }
}

template <typename T>
struct MfmaAcc;
// Operands: lane (c, g) = (lane % 16, lane / 16) holds A[c][g] and B[g][c].  Result element r
// of the lane is D[row(g, r)][c]: row 4g + r for FP32, 4r + g for FP64 (measured on gfx950).
template <>
struct MfmaAcc<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * r + g; }
};
template <>
struct MfmaAcc<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// y[node[t]] += sum of part[ent[off[t] .. off[t+1])] * bs components: the skeleton nodes'
// element contributions, in slot order.
// With cmask: components flagged in cmask[t] are Dirichlet rows, set to cx (the constrained
// operator's y = x there), or to zero without cx (a restriction onto constrained dofs), instead.
// With a Chebyshev step (cs.p != null, overwriting apply): the node's product t feeds the step
// (p and x updated there, y not written).
template <typename T, int BS>
__global__ void skeleton_gather_kernel(const ptrdiff_t n, const int32_t *__restrict__ node,
                                       const int32_t *__restrict__ off, const int32_t *__restrict__ ent,
                                       const T *__restrict__ part, T *__restrict__ y, const int overwrite,
                                       const uint8_t *__restrict__ cmask, const T *cx, const ChebStep<T> cs) {
    // the first NU partials of a node (face nodes have 2, edge nodes 4) are loaded together with its
    // old value, right after its offsets: two memory round trips per node instead of one per partial
    constexpr int NU = 4;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const int j0 = off[t], j1 = off[t + 1];
        const ptrdiff_t g = (ptrdiff_t)node[t] * BS;
        const unsigned cm = cmask ? cmask[t] : 0u;
        T acc[BS], yo[BS];
#pragma unroll
        for (int c = 0; c < BS; c++) acc[c] = yo[c] = 0;
        if (!ent) {
            // absent slots re-read the node's first partial (a valid index: every skeleton node has
            // one) and are zeroed after the load, so that no load waits behind a branch
            T pv[NU][BS];
#pragma unroll
            for (int k = 0; k < NU; k++) {
                const ptrdiff_t q = (ptrdiff_t)(j0 + k < j1 ? j0 + k : j0) * BS;
#pragma unroll
                for (int c = 0; c < BS; c++) pv[k][c] = part[q + c];
            }
            if (!overwrite)
#pragma unroll
                for (int c = 0; c < BS; c++) yo[c] = y[g + c];
            // in slot order, as the loop below (adding the zeros of absent slots leaves the sum's
            // words unchanged: it starts at +0 and never becomes -0)
#pragma unroll
            for (int k = 0; k < NU; k++)
#pragma unroll
                for (int c = 0; c < BS; c++) acc[c] += j0 + k < j1 ? pv[k][c] : T(0);
            for (int j = j0 + NU; j < j1; j++)
#pragma unroll
                for (int c = 0; c < BS; c++) acc[c] += part[(ptrdiff_t)j * BS + c];
        } else {
            if (!overwrite)
#pragma unroll
                for (int c = 0; c < BS; c++) yo[c] = y[g + c];
            for (int j = j0; j < j1; j++) {
                const ptrdiff_t q = (ptrdiff_t)ent[j] * BS;
#pragma unroll
                for (int c = 0; c < BS; c++) acc[c] += part[q + c];
            }
        }
#pragma unroll
        for (int c = 0; c < BS; c++) {
            const T v = ((cm >> c) & 1u) ? (cx ? cx[g + c] : T(0)) : (overwrite ? acc[c] : yo[c] + acc[c]);
            if (cs.p) {
                const T pi = cheb3_direction<T>(cs.first ? T(0) : cs.p[g + c], cs.rhs[g + c], v, cs.beta, cs.dinv, g + c,
                                                cs.first);
                cs.p[g + c] = pi;
                cs.x[g + c] = __builtin_fma(-cs.alpha, pi, cs.x[g + c]);
            } else {
                y[g + c] = v;
            }
        }
    }
}

// d += diag(A): node-wise sum of the 8x8 diagonal over the micro elements present
// (operators/sshex8/sshex8_laplacian.cpp:333-435).
template <typename T, typename G, int L>
__global__ void __launch_bounds__(LapShape<L>::BLOCK)
    sshex8_laplacian_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx, const G *__restrict__ fff,
                                 const ptrdiff_t fstride, const float *__restrict__ coeff,
                                 const double *__restrict__ Kem, T *__restrict__ d) {
    using S = LapShape<L>;
    constexpr int N1 = S::N1, NXE = S::NXE, LINES = S::LINES, EPB = S::EPB;
    __shared__ double sK[EPB][64];
    const int tid = threadIdx.x;
    const int le = tid / LINES;
    const int line = tid - le * LINES;
    const int yy = line % N1, zz = line / N1;
    const ptrdiff_t e0 = (ptrdiff_t)blockIdx.x * EPB;
    const ptrdiff_t e = e0 + le;
    const bool active = (le < EPB) && (e < E);
    lap_block_matrices<L, EPB, S::BLOCK, G>(E, e0, fff, fstride, coeff, Kem, sK);
    __syncthreads();
    if (!active) return;
    const double *K = sK[le];
    for (int i = 0; i <= L; i++) {
        double acc = 0;
        for (int c = 0; c < 8; c++) {
            const int cx = hex8_cx(c), cy = hex8_cy(c), cz = hex8_cz(c);
            const bool ok = (cx ? i > 0 : i < L) && (cy ? yy > 0 : yy < L) && (cz ? zz > 0 : zz < L);
            if (ok) acc += K[c * 9];
        }
        atomic_add(&d[idx[e * NXE + lidx(L, i, yy, zz)]], (T)acc);
    }
}

// ------------------------------------------------------------------ elasticity
// Per-macro 24x24 sub-element matrix (index component*8 + node), computed once per
// operator setup: adj * h^2, det * h^3 (operators/sshex8/sshex8.hpp:26-42), mu, lambda
// scaled by the optional coefficient.
template <typename G>
__global__ void sshex8_elasticity_matrices_kernel(const ptrdiff_t E, const int L, const G *__restrict__ adj,
                                                  const G *__restrict__ det, const ptrdiff_t gstride,
                                                  const double mu, const double lambda,
                                                  const float *__restrict__ coeff, double *__restrict__ Kout) {
    const ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * 576) return;
    const ptrdiff_t e = t / 576;
    const int rc = (int)(t - e * 576);
    const double h = 1.0 / L;
    double a[9];
#pragma unroll
    for (int q = 0; q < 9; q++) a[q] = load_geom<G>(adj, q * gstride + e) * (h * h);
    const double dt = load_geom<G>(det, e) * (h * h * h);
    const double c = coeff ? (double)coeff[e] : 1.0;
    Kout[t] = hex8_elasticity_entry(a, dt, mu * c, lambda * c, rc / 24, rc % 24);
}

template <int L>
struct ElShape {
    static constexpr int N1 = L + 1;
    static constexpr int PLANE = N1 * N1;
    // keep the per-block copies of the 24x24 matrices small in LDS
    static constexpr int EPB = PLANE >= 64 ? (PLANE >= 256 ? 1 : 256 / PLANE) : 4;
    static constexpr int BLOCK = round64(EPB * PLANE);
};

// y += A x for sshex8 linear elasticity (x, y interleaved xyz).  One block per macro
// element (or EPB small ones); the element is swept plane by plane in z through a ring of
// three LDS planes (2.5-D streaming, so L = 16 fits), one thread per node of the plane.
// Each node sums over the (up to 8) micro elements that contain it, 3 rows x 24 columns
// of the per-macro matrix each (K rows broadcast from LDS).  Interior nodes update y (their
// y is loaded with the next plane of x); skeleton nodes store their three components to the
// element's own slots of `part` (element-major: 24-byte runs), summed per node by
// skeleton_gather_kernel through the node-sorted slot list (no atomics).
template <typename T, int L>
__global__ void __launch_bounds__(ElShape<L>::BLOCK)
    sshex8_elasticity_apply_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                   const T *__restrict__ Kall, const T *__restrict__ x,
                                   T *__restrict__ y, T *__restrict__ part, const int overwrite) {
    using S = ElShape<L>;
    constexpr int N1 = S::N1, NXE = N1 * N1 * N1, PLANE = S::PLANE, EPB = S::EPB, NSK = sk_count(L);
    __shared__ T su[EPB][3][PLANE][3];
    // one element per block: its matrix is block-uniform and is read straight from memory into
    // scalar registers; several elements per block: staged in LDS
    __shared__ T sK[EPB == 1 ? 1 : EPB][EPB == 1 ? 1 : 576];

    const int tid = threadIdx.x;
    const int le = tid / PLANE;
    const int p = tid - le * PLANE;
    const int xx = p % N1, yy = p / N1;
    const ptrdiff_t e0 = (ptrdiff_t)blockIdx.x * EPB;
    const ptrdiff_t e = e0 + le;
    const bool active = (le < EPB) && (e < E);
    const bool edge_xy = xx == 0 || xx == L || yy == 0 || yy == L;

    if constexpr (EPB > 1) {
        for (int k = tid; k < EPB * 576; k += S::BLOCK) {
            const int l2 = k / 576;
            const ptrdiff_t e2 = e0 + l2;
            sK[l2][k - l2 * 576] = (e2 < E) ? Kall[e2 * 576 + (k - l2 * 576)] : T(0);
        }
    }
    const int32_t *row = idx + (active ? e : 0) * NXE;
    int32_t g_next = 0;
    auto load_plane = [&](int z) {
        if (active) {
            g_next = row[lidx(L, xx, yy, z)];
#pragma unroll
            for (int c = 0; c < 3; c++) su[le][z % 3][p][c] = x[(ptrdiff_t)g_next * 3 + c];
        }
    };
    load_plane(0);
    __syncthreads();
    T *pe = part + e * (ptrdiff_t)NSK * 3;  // the element's skeleton slots (element-major)

    for (int z = 0; z <= L; z++) {
        const int32_t g = g_next;
        const bool interior = !edge_xy && z > 0 && z < L;
        T yold[3] = {0, 0, 0};
        if (active && interior && !overwrite) {
#pragma unroll
            for (int c = 0; c < 3; c++) yold[c] = y[(ptrdiff_t)g * 3 + c];
        }
        if (z + 1 <= L) {
            // slot (z+1)%3 held plane z-2, which nobody reads any more
            load_plane(z + 1);
        }
        __syncthreads();
        if (active) {
            T acc[3] = {0, 0, 0};
            const T *K = EPB == 1 ? Kall + e0 * 576 : &sK[le][0];
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int cx = hex8_cx(c), cy = hex8_cy(c), cz = hex8_cz(c);
                const bool ok = (cx ? xx > 0 : xx < L) && (cy ? yy > 0 : yy < L) && (cz ? z > 0 : z < L);
                if (ok) {
                    // micro element with lower corner (xx-cx, yy-cy, z-cz); this node is its corner c
                    T ue[24];
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int nx = xx - cx + hex8_cx(j), ny = yy - cy + hex8_cy(j), nz = z - cz + hex8_cz(j);
                        const T *src = su[le][nz % 3][nx + ny * N1];
                        ue[0 * 8 + j] = src[0];
                        ue[1 * 8 + j] = src[1];
                        ue[2 * 8 + j] = src[2];
                    }
#pragma unroll
                    for (int ci = 0; ci < 3; ci++) {
                        const T *Kr = K + (ci * 8 + c) * 24;
                        T s = 0;
#pragma unroll
                        for (int j = 0; j < 24; j++) s += Kr[j] * ue[j];
                        acc[ci] += s;
                    }
                }
            }
            if (interior) {
#pragma unroll
                for (int ci = 0; ci < 3; ci++) y[(ptrdiff_t)g * 3 + ci] = yold[ci] + acc[ci];
            } else {
                T *q = pe + (ptrdiff_t)sk_slot(L, xx, yy, z) * 3;
#pragma unroll
                for (int ci = 0; ci < 3; ci++) q[ci] = acc[ci];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// sshex8 elasticity in stencil form.  All micro elements of an affine macro element share one
// 24x24 matrix K, so at a node whose 8 surrounding micro elements lie in the element the operator
// is one 27-point stencil of 3x3 blocks (243 coefficients).  Split by the z offset d of the
// neighbour: S_m (d = -1), S_p (d = +1) and the in-plane part S_0 = S_0lo + S_0up, where S_0lo
// collects the micro elements above the node and S_0up those below (a node on the z = 0 face sees
// only S_0lo, on z = L only S_0up).  Coefficient order within a block: [nbr 9][i][j],
// nbr = (dy+1)*3 + dx+1, i the output component, j the input component.
__host__ __device__ constexpr inline int st_nbr(int dx, int dy) { return (dy + 1) * 3 + dx + 1; }

// ------------------------------------------------------------------------------------------
// FP32 sshex8 elasticity on the matrix cores (the operator of the FP32 multigrid preconditioner).
// Element-by-element: all L^3 micro elements of a macro element share its 24x24 matrix K, so
// Y = K X with X = the micro elements' gathered corner values (24 x L^3) is one batched GEMM.
// It runs as v_mfma_f32_16x16x32_bf16 tiles: M = the 24 rows of K (two 16-row tiles), K = its 24
// columns (8 corners x 3 components, padded to 32), N = 16 micro elements.  FP32 operands are
// split into bf16 hi + lo parts and multiplied as hi*hi + hi*lo + lo*hi (~2^-17 relative per
// product, FP32 accumulation), far below the BF16-geometry error of this operator.  The micro
// elements are visited in 8 parity colours (lower corner parity in x, y, z): two micro elements
// of one colour share no node, so each colour's contributions are added to the LDS node
// accumulators without atomics, colour by colour (a fixed order: deterministic).  Interior nodes
// then update y; skeleton nodes store to the element's slots of `part` (as the vector kernel).
// Skeleton slot (sk_slot) of each local node id, -1 for interior nodes: one table load in place
// of the coordinate division and the face/edge tests per node.
// LDS layout of the FP32 matrix-core kernel's node accumulators: node (x, y, z) of component
// plane c at c * PS + z * PL + y * RS + x.  LDS serves a half-wave (32 lanes) per pass and dword
// slot from 32 banks; with the padded L = 8 strides the read-modify-writes of one parity colour
// (the corners of 16 micro elements in two component planes per half-wave) meet each bank once
// per pass, against twice with the dense strides (bank model checked against SQ_LDS_BANK_CONFLICT).
template <int L>
struct ElLds {
    static constexpr int RS = L + 1, PL = (L + 1) * (L + 1), PS = (L + 1) * (L + 1) * (L + 1);
};
template <>
struct ElLds<8> {
    static constexpr int RS = 12, PL = 109, PS = 977;
};

// The FP64 kernel's node planes (x and accumulators, 8-byte words).  Same bank model: with the
// padded L = 8 strides the B-operand reads and accumulator updates of a parity colour need half
// the passes of the dense layout.
template <int L>
struct ElLds64 {
    static constexpr int RS = L + 1, PL = (L + 1) * (L + 1), PS = (L + 1) * (L + 1) * (L + 1);
};
template <>
struct ElLds64<8> {
    static constexpr int RS = 10, PL = 90, PS = 809;
};

template <int L, typename Lay>
struct SkSlotTable {
    int16_t v[(L + 1) * (L + 1) * (L + 1)];    // skeleton slot
    int16_t lds[(L + 1) * (L + 1) * (L + 1)];  // Lay position in a component plane
    constexpr SkSlotTable() : v(), lds() {
        for (int z = 0; z <= L; z++)
            for (int y = 0; y <= L; y++)
                for (int x = 0; x <= L; x++) {
                    v[lidx(L, x, y, z)] = (int16_t)sk_slot(L, x, y, z);
                    lds[lidx(L, x, y, z)] = (int16_t)(z * Lay::PL + y * Lay::RS + x);
                }
    }
};
template <int L, typename Lay = ElLds<L>>
__constant__ SkSlotTable<L, Lay> sk_slot_table = SkSlotTable<L, Lay>();

template <int L>
struct ElMfma {
    static constexpr int N1 = L + 1, NXE = N1 * N1 * N1, H = L / 2;
    static constexpr int NC = H * H * H;                 // micro elements per colour
    static constexpr int TPC = (NC + 15) / 16;           // 16-column tiles per colour
    static constexpr int EPB = TPC >= 4 ? 1 : 4 / TPC;   // macro elements per 256-thread block
    static constexpr int WPE = 4 / EPB;                  // waves per macro element
};

typedef __bf16 el_b8 __attribute__((ext_vector_type(8)));
typedef float el_f4 __attribute__((ext_vector_type(4)));

// v = hi + lo (+ ~2^-17 |v|), both bf16 (round to nearest even, v_cvt_pk_bf16_f32), packed as
// hi | lo << 16
__device__ __forceinline__ uint32_t el_split_pack(float v) {
    const __bf16 h = (__bf16)v;
    const __bf16 l = (__bf16)(v - (float)h);
    return (uint32_t)__builtin_bit_cast(uint16_t, h) | ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
}
// eight packed pairs -> the hi and lo bf16x8 operands
__device__ __forceinline__ void el_unpack8(const uint32_t (&p)[8], el_b8 &hi, el_b8 &lo) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    u4 h, l;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        h[i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x05040100u);
        l[i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x07060302u);
    }
    hi = __builtin_bit_cast(el_b8, h);
    lo = __builtin_bit_cast(el_b8, l);
}

typedef uint32_t el_u4 __attribute__((ext_vector_type(4)));

// The A operands of the FP32 kernel below, split once per matrix: for matrix e, 16-row tile tt and
// lane (c, g) = (lane % 16, lane / 16), the bf16 hi and lo parts of K rows 16 tt + c, columns
// 8 g .. 8 g + 7 (zero outside the 24 x 24 matrix), at Apk[((e * 2 + tt) * 2 + part) * 64 + lane].
__global__ void sshex8_elasticity_pack_bf16x2_kernel(const ptrdiff_t nK, const float *__restrict__ K,
                                                     el_u4 *__restrict__ Apk) {
    const ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nK * 128) return;
    const ptrdiff_t e = t >> 7;
    const int tt = (int)(t >> 6) & 1, lane = (int)t & 63, c = lane & 15, g = lane >> 4;
    const int row = 16 * tt + c;
    const bool ok = row < 24 && g < 3;
    const float *kr = K + e * 576 + (ok ? row * 24 + 8 * g : 0);
    uint32_t pk[8];
#pragma unroll
    for (int j = 0; j < 8; j++) pk[j] = el_split_pack(ok ? kr[j] : 0.f);
    el_b8 hi, lo;
    el_unpack8(pk, hi, lo);
    Apk[((e * 2 + tt) * 2 + 0) * 64 + lane] = __builtin_bit_cast(el_u4, hi);
    Apk[((e * 2 + tt) * 2 + 1) * 64 + lane] = __builtin_bit_cast(el_u4, lo);
}

int sshex8_elasticity_pack_bf16x2(ptrdiff_t nK, const float *K, uint32_t *Apk, hipStream_t s) {
    if (nK == 0) return SFEM_SUCCESS;
    sshex8_elasticity_pack_bf16x2_kernel<<<(unsigned)((nK * 128 + 255) / 256), 256, 0, s>>>(nK, K, (el_u4 *)Apk);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

// Consecutive groups of 2^kshift elements share one matrix (kshift > 0: the sub-blocks of a macro
// element of a higher level, see sshex8_subblock_idx).  Apk: the matrices' packed A operands
// (sshex8_elasticity_pack_bf16x2).
template <int L>
__global__ void __launch_bounds__(256)
    sshex8_elasticity_apply_mfma_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                        const el_u4 *__restrict__ Apk, const float *x, float *__restrict__ y,
                                        float *__restrict__ part, const int overwrite, const int kshift,
                                        const int32_t *__restrict__ sk_pos, const ChebStep<float> cs) {
    using S = ElMfma<L>;
    constexpr int N1 = S::N1, NXE = S::NXE, H = S::H, NC = S::NC, TPC = S::TPC, EPB = S::EPB, WPE = S::WPE;
    constexpr int NSK = sk_count(L);
    constexpr int NPT = (EPB * NXE + 255) / 256;  // nodes per thread in the load and store phases
    using Y = ElLds<L>;
    __shared__ uint32_t sx[EPB][3][NXE];  // x split into bf16 hi/lo pairs, one plane per component
    __shared__ float sy[EPB][3][Y::PS];   // node accumulators

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const ptrdiff_t e0 = xcd_contiguous_block(blockIdx.x, gridDim.x) * EPB;

    // node k = tid + 256 i of the block: ids first, then all values at once; interior nodes
    // start from their y (accumulate), skeleton nodes from 0
    int32_t gid[NPT], pos[NPT];  // node ids; node-sorted skeleton slots (-1: interior)
    int lds[NPT];                // LDS positions
    float xk[NPT][3];            // x of the node (the Chebyshev step updates interior ones)
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE;
        gid[i] = (k < EPB * NXE && e0 + le < E) ? idx[e0 * NXE + k] : -1;
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE, v = k - le * NXE;
        const int slot = sk_slot_table<L>.v[v];
        lds[i] = sk_slot_table<L>.lds[v];
        const bool interior = slot < 0;
        pos[i] = (gid[i] >= 0 && !interior) ? sk_pos[(e0 + le) * NSK + slot] : -1;
        float xv[3] = {0, 0, 0}, yv[3] = {0, 0, 0};
        if (gid[i] >= 0) {
            const float *xp = x + (ptrdiff_t)gid[i] * 3;
            const float *yp = y + (ptrdiff_t)gid[i] * 3;
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                xv[ci] = xp[ci];
                if (interior && !overwrite) yv[ci] = yp[ci];
            }
        }
        if (k < EPB * NXE) {
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                sx[le][ci][v] = el_split_pack(xv[ci]);
                sy[le][ci][lds[i]] = yv[ci];
            }
        }
#pragma unroll
        for (int ci = 0; ci < 3; ci++) xk[i][ci] = xv[ci];
    }

    // A operands (this wave's macro element): tile tt holds K rows 16tt + c, columns 8g + j
    const int le = wave / WPE, wsub = wave - le * WPE;
    const ptrdiff_t e = e0 + le;
    const bool active = e < E;
    el_b8 ahi[2], alo[2];
    {
        const el_u4 *ak = Apk + (active ? (e >> kshift) : 0) * 256 + lane;
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
            ahi[tt] = __builtin_bit_cast(el_b8, ak[(2 * tt) * 64]);
            alo[tt] = __builtin_bit_cast(el_b8, ak[(2 * tt + 1) * 64]);
        }
    }
    __syncthreads();

    // Tiles of 16 micro elements of one colour: wave wsub takes tiles wsub, wsub + WPE, ...  Lane
    // (c, g) feeds B rows 8g .. 8g + 7 of column c (component g, corners 0..7; lanes g = 3 meet zero
    // columns of A and read plane 2 with the g = 2 lanes of their half-wave, adding no bank traffic)
    // and receives D rows 16tt + 4g + r: component (16tt + 4g) / 8, corner r (g even) or r + 4 (g
    // odd, one plane up in z).  Addresses: per lane and tile for colour 0, plus a wave-uniform
    // colour shift.
    constexpr int TPW = (TPC + WPE - 1) / WPE;
    const int gc = g < 3 ? g : 2;
    const uint32_t *sxl = &sx[0][0][0] + (le * 3 + gc) * NXE;
    float *syl[2];
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        const int ci = (16 * tt + 4 * g) >> 3;
        syl[tt] = &sy[0][0][0] + (le * 3 + (ci < 3 ? ci : 0)) * Y::PS + ((g & 1) ? Y::PL : 0);
    }
    const bool row_ok[2] = {true, g < 2};
    int xb[TPW], yb[TPW];
    bool valid[TPW];
#pragma unroll
    for (int u = 0; u < TPW; u++) {
        const int t = wsub + u * WPE;
        const unsigned q = 16 * t + c;
        valid[u] = active && t < TPC && q < (unsigned)NC;
        const unsigned qq = valid[u] ? q : 0;
        const int mx = 2 * (qq % H), my = 2 * ((qq / H) % H), mz = 2 * (qq / (H * H));
        xb[u] = lidx(L, mx, my, mz);
        yb[u] = mz * Y::PL + my * Y::RS + mx;
    }
    // B columns of tile u in colour `colour` (in range for every lane: qq, gc clamped; invalid
    // columns are never scattered)
    auto load_b = [&](int colour, int u, uint32_t(&pk)[8]) {
        const uint32_t *xp = sxl + xb[u] + lidx(L, colour & 1, (colour >> 1) & 1, colour >> 2);
#pragma unroll
        for (int j = 0; j < 8; j++) pk[j] = xp[lidx(L, hex8_cx(j), hex8_cy(j), hex8_cz(j))];
    };
    // With one tile per wave and colour the next colour's columns are read from LDS before this
    // colour's products (x in LDS does not change in the loop), and each tile's node accumulators
    // before its products (the two row tiles hold different components, and no other wave touches
    // these nodes in this colour): the LDS latency overlaps the matrix-core work
    uint32_t pkn[8];
    if (TPW == 1 && wsub < TPC) load_b(0, 0, pkn);
#pragma unroll 1
    for (int colour = 0; colour < 8; colour++) {
        const int px = colour & 1, py = (colour >> 1) & 1, pz = colour >> 2;
        const int yc = pz * Y::PL + py * Y::RS + px;
#pragma unroll
        for (int u = 0; u < TPW; u++) {
            if (wsub + u * WPE >= TPC) break;  // wave-uniform
            uint32_t pk[8];
            if (TPW == 1) {
#pragma unroll
                for (int j = 0; j < 8; j++) pk[j] = pkn[j];
                if (colour < 7) load_b(colour + 1, 0, pkn);
            } else {
                load_b(colour, u, pk);
            }
            float yv[2][4];
#pragma unroll
            for (int tt = 0; tt < 2; tt++)
                if (valid[u] && row_ok[tt]) {
                    const float *yp = syl[tt] + yb[u] + yc;
                    yv[tt][0] = yp[0];
                    yv[tt][1] = yp[1];
                    yv[tt][2] = yp[Y::RS + 1];
                    yv[tt][3] = yp[Y::RS];
                }
            el_b8 bhi, blo;
            el_unpack8(pk, bhi, blo);
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                el_f4 d = {0.f, 0.f, 0.f, 0.f};
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], bhi, d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], blo, d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[tt], bhi, d, 0, 0, 0);
                if (valid[u] && row_ok[tt]) {
                    float *yp = syl[tt] + yb[u] + yc;
                    yp[0] = yv[tt][0] + d[0];
                    yp[1] = yv[tt][1] + d[1];
                    yp[Y::RS + 1] = yv[tt][2] + d[2];
                    yp[Y::RS] = yv[tt][3] + d[3];
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le2 = k / NXE, v = k - le2 * NXE;
        if (gid[i] < 0) continue;
        if (pos[i] < 0 && cs.p) {  // interior node: its product is complete here
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                const ptrdiff_t q = (ptrdiff_t)gid[i] * 3 + ci;
                const float pi = cheb3_direction<float>(cs.first ? 0.f : cs.p[q], cs.rhs[q], sy[le2][ci][lds[i]],
                                                        cs.beta, cs.dinv, q, cs.first);
                cs.p[q] = pi;
                cs.x[q] = __builtin_fma(-cs.alpha, pi, xk[i][ci]);
            }
            continue;
        }
        float *q = pos[i] < 0 ? y + (ptrdiff_t)gid[i] * 3 : part + (ptrdiff_t)pos[i] * 3;
#pragma unroll
        for (int ci = 0; ci < 3; ci++) q[ci] = sy[le2][ci][lds[i]];
    }
}

// FP64 sshex8 elasticity on the matrix cores: the same element-by-element GEMM as the FP32 kernel
// above, in exact FP64 with v_mfma_f64_16x16x4f64: M = the 24 rows of K (two 16-row tiles, rows
// 24..31 zero), K = its 24 columns (six steps of 4), N = 16 micro elements of one parity colour.
// The node sums run colour by colour (deterministic); they differ from the node-wise vector kernel
// only in the order of the FP64 additions.
template <int L>
__global__ void __launch_bounds__(256)
    sshex8_elasticity_apply_mfma64_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                          const double *__restrict__ Kall, const double *x, double *__restrict__ y,
                                          double *__restrict__ part, const int overwrite, const int kshift,
                                          const int32_t *__restrict__ sk_pos, const ChebStep<double> cs) {
    using S = ElMfma<L>;
    using M = MfmaAcc<double>;
    using Y = ElLds64<L>;
    constexpr int N1 = S::N1, NXE = S::NXE, H = S::H, NC = S::NC, TPC = S::TPC, EPB = S::EPB, WPE = S::WPE;
    constexpr int NSK = sk_count(L);
    constexpr int NPT = (EPB * NXE + 255) / 256;
    __shared__ double sx[EPB][3][Y::PS];
    __shared__ double sy[EPB][3][Y::PS];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const ptrdiff_t e0 = xcd_contiguous_block(blockIdx.x, gridDim.x) * EPB;

    int32_t gid[NPT], pos[NPT];  // node ids; node-sorted skeleton slots (-1: interior)
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE;
        gid[i] = (k < EPB * NXE && e0 + le < E) ? idx[e0 * NXE + k] : -1;
    }
    int lds[NPT];  // LDS positions
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE, v = k - le * NXE;
        const int slot = sk_slot_table<L, Y>.v[v];
        lds[i] = sk_slot_table<L, Y>.lds[v];
        const bool interior = slot < 0;
        pos[i] = (gid[i] >= 0 && !interior) ? sk_pos[(e0 + le) * NSK + slot] : -1;
        double xv[3] = {0, 0, 0}, yv[3] = {0, 0, 0};
        if (gid[i] >= 0) {
            const double *xp = x + (ptrdiff_t)gid[i] * 3;
            const double *yp = y + (ptrdiff_t)gid[i] * 3;
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                xv[ci] = xp[ci];
                if (interior && !overwrite) yv[ci] = yp[ci];
            }
        }
        if (k < EPB * NXE) {
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                sx[le][ci][lds[i]] = xv[ci];
                sy[le][ci][lds[i]] = yv[ci];
            }
        }
    }

    // A: tile tt, step s holds K[16tt + c][4s + g]
    const int le = wave / WPE, wsub = wave - le * WPE;
    const ptrdiff_t e = e0 + le;
    const bool active = e < E;
    double a[2][6];
#pragma unroll
    for (int tt = 0; tt < 2; tt++)
#pragma unroll
        for (int s = 0; s < 6; s++) {
            const int row = 16 * tt + c;
            a[tt][s] = (active && row < 24) ? Kall[(e >> kshift) * 576 + row * 24 + 4 * s + g] : 0.0;
        }
    __syncthreads();

#pragma unroll 1
    for (int colour = 0; colour < 8; colour++) {
        const int px = colour & 1, py = (colour >> 1) & 1, pz = colour >> 2;
        for (int t = wsub; t < TPC; t += WPE) {
            const int q = 16 * t + c;
            const bool valid = active && q < NC;
            const int qq = valid ? q : 0;
            const int mx = 2 * (qq % H) + px, my = 2 * ((qq / H) % H) + py, mz = 2 * (qq / (H * H)) + pz;
            // the lane's accumulators (distinct (component, corner) rows: distinct nodes or planes,
            // untouched by other waves in this colour), read before the products so that the LDS
            // latency overlaps them
            double yv[4][2];
            if (valid) {
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int tt = 0; tt < 2; tt++) {
                        const int R = 16 * tt + M::row(g, r);  // component R / 8, corner R % 8
                        if (R < 24) {
                            const int k = R & 7;
                            yv[r][tt] = sy[le][R >> 3][(mz + hex8_cz(k)) * Y::PL + (my + hex8_cy(k)) * Y::RS + mx +
                                                      hex8_cx(k)];
                        }
                    }
            }
            typename M::v4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 6; s++) {
                const int kk = 4 * s + g, j = kk & 7;  // column: component kk / 8, corner kk % 8
                const int node = (mz + hex8_cz(j)) * Y::PL + (my + hex8_cy(j)) * Y::RS + mx + hex8_cx(j);
                const double b = sx[le][kk >> 3][node];  // in range (qq clamped); invalid columns unused
                d0 = M::step(a[0][s], b, d0);
                d1 = M::step(a[1][s], b, d1);
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
#pragma unroll
                    for (int tt = 0; tt < 2; tt++) {
                        const int R = 16 * tt + M::row(g, r);
                        if (R < 24) {
                            const int k = R & 7;
                            sy[le][R >> 3][(mz + hex8_cz(k)) * Y::PL + (my + hex8_cy(k)) * Y::RS + mx + hex8_cx(k)] =
                                    yv[r][tt] + (tt ? d1[r] : d0[r]);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le2 = k / NXE, v = k - le2 * NXE;
        if (gid[i] < 0) continue;
        if (pos[i] < 0 && cs.p) {  // interior node: its product is complete here
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                const ptrdiff_t q = (ptrdiff_t)gid[i] * 3 + ci;
                const double pi = cheb3_direction<double>(cs.first ? 0.0 : cs.p[q], cs.rhs[q], sy[le2][ci][lds[i]],
                                                          cs.beta, cs.dinv, q, cs.first);
                cs.p[q] = pi;
                cs.x[q] = __builtin_fma(-cs.alpha, pi, sx[le2][ci][lds[i]]);
            }
            continue;
        }
        double *q = pos[i] < 0 ? y + (ptrdiff_t)gid[i] * 3 : part + (ptrdiff_t)pos[i] * 3;
#pragma unroll
        for (int ci = 0; ci < 3; ci++) q[ci] = sy[le2][ci][lds[i]];
    }
}

// Small levels (L = 2: 8 micro elements): the colour-by-colour scatter above would fill 1/16 of
// each 16-column tile and pay a barrier per colour.  Here one wave takes one macro
// element and runs all of its micro elements through the GEMM at once (one or four tiles), the 24
// results of each micro element go to an LDS table Y[micro][row], and each node then sums its
// (at most 8) entries in a fixed order (z, y, x of the micro element ascending): deterministic, no
// atomics.  FP32: bf16x3 products as in the colour kernel (same packed A operands); FP64:
// v_mfma_f64_16x16x4f64.  Interior nodes are stored (or fed to a Chebyshev step), skeleton nodes go
// to their node-sorted partial slots.
template <typename T, int L>
struct ElSmall {
    static constexpr int N1 = L + 1, NXE = N1 * N1 * N1, NM = L * L * L, TILES = (NM + 15) / 16;
    static constexpr int EPB = 4;                                  // one wave per macro element
    static constexpr int NPT = (EPB * NXE + 255) / 256;            // nodes per thread
    static constexpr int XW = std::is_same<T, double>::value ? 2 : 1;  // x words per value in LDS
};

template <typename T, int L>
__global__ void __launch_bounds__(256)
    sshex8_elasticity_apply_small_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx, const T *__restrict__ Kall,
                                         const el_u4 *__restrict__ Apk, const T *x, T *__restrict__ y,
                                         T *__restrict__ part, const int overwrite, const int kshift,
                                         const int32_t *__restrict__ sk_pos, const ChebStep<T> cs) {
    using S = ElSmall<T, L>;
    constexpr int N1 = S::N1, NXE = S::NXE, NM = S::NM, EPB = S::EPB, NPT = S::NPT;
    constexpr int NSK = sk_count(L);
    constexpr bool F64 = std::is_same<T, double>::value;
    typedef typename std::conditional<F64, double, uint32_t>::type XWord;  // FP32: bf16 hi | lo
    __shared__ XWord sx[EPB][3][NXE];
    __shared__ T sY[EPB][NM][24];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c = lane & 15, g = lane >> 4;
    const ptrdiff_t e0 = (ptrdiff_t)blockIdx.x * EPB;

    int32_t gid[NPT], pos[NPT];
    T xk[NPT][3], yk[NPT][3];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE;
        gid[i] = (k < EPB * NXE && e0 + le < E) ? idx[e0 * NXE + k] : -1;
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le = k / NXE, v = k - le * NXE;
        const int slot = sk_slot_table<L>.v[v];
        const bool interior = slot < 0;
        pos[i] = (gid[i] >= 0 && !interior) ? sk_pos[(e0 + le) * NSK + slot] : -1;
#pragma unroll
        for (int ci = 0; ci < 3; ci++) {
            xk[i][ci] = gid[i] >= 0 ? x[(ptrdiff_t)gid[i] * 3 + ci] : T(0);
            yk[i][ci] = (gid[i] >= 0 && interior && !overwrite && !cs.p) ? y[(ptrdiff_t)gid[i] * 3 + ci] : T(0);
        }
        if (k < EPB * NXE) {
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                if constexpr (F64)
                    sx[le][ci][v] = xk[i][ci];
                else
                    sx[le][ci][v] = el_split_pack(xk[i][ci]);
            }
        }
    }

    const int le = wave;
    const ptrdiff_t e = e0 + le;
    const bool active = e < E;
    const ptrdiff_t ek = active ? (e >> kshift) : 0;
    // A operands of the element's matrix
    el_b8 ahi[2], alo[2];
    double a64[2][6];
    if constexpr (F64) {
#pragma unroll
        for (int tt = 0; tt < 2; tt++)
#pragma unroll
            for (int st = 0; st < 6; st++) {
                const int row = 16 * tt + c;
                a64[tt][st] = (active && row < 24) ? (double)Kall[ek * 576 + row * 24 + 4 * st + g] : 0.0;
            }
    } else {
        const el_u4 *ak = Apk + ek * 256 + lane;
#pragma unroll
        for (int tt = 0; tt < 2; tt++) {
            ahi[tt] = __builtin_bit_cast(el_b8, ak[(2 * tt) * 64]);
            alo[tt] = __builtin_bit_cast(el_b8, ak[(2 * tt + 1) * 64]);
        }
    }
    __syncthreads();

#pragma unroll
    for (int t = 0; t < S::TILES; t++) {
        const int q = 16 * t + c;
        const bool valid = active && q < NM;
        const int qq = valid ? q : 0;
        const int mx = qq % L, my = (qq / L) % L, mz = qq / (L * L);
        const int base = lidx(L, mx, my, mz);
        if constexpr (F64) {
            using M = MfmaAcc<double>;
            typename M::v4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
#pragma unroll
            for (int st = 0; st < 6; st++) {
                const int kk = 4 * st + g, j = kk & 7;  // column: component kk / 8, corner kk % 8
                const double b = sx[le][kk >> 3][base + lidx(L, hex8_cx(j), hex8_cy(j), hex8_cz(j))];
                d0 = M::step(a64[0][st], b, d0);
                d1 = M::step(a64[1][st], b, d1);
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int R0 = M::row(g, r), R1 = 16 + M::row(g, r);
                    sY[le][q][R0] = d0[r];
                    if (R1 < 24) sY[le][q][R1] = d1[r];
                }
            }
        } else {
            const int gc = g < 3 ? g : 2;  // lanes g = 3 meet zero columns of A
            uint32_t pk[8];
#pragma unroll
            for (int j = 0; j < 8; j++) pk[j] = sx[le][gc][base + lidx(L, hex8_cx(j), hex8_cy(j), hex8_cz(j))];
            el_b8 bhi, blo;
            el_unpack8(pk, bhi, blo);
#pragma unroll
            for (int tt = 0; tt < 2; tt++) {
                el_f4 d = {0.f, 0.f, 0.f, 0.f};
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], bhi, d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], blo, d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[tt], bhi, d, 0, 0, 0);
                const int R0 = 16 * tt + 4 * g;  // rows R0 .. R0 + 3
                if (valid && R0 < 24) {
#pragma unroll
                    for (int r = 0; r < 4; r++) sY[le][q][R0 + r] = d[r];
                }
            }
        }
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int k = tid + 256 * i, le2 = k / NXE, v = k - le2 * NXE;
        if (gid[i] < 0) continue;
        const int xx = v % N1, yy = (v / N1) % N1, zz = v / (N1 * N1);
        T acc[3] = {0, 0, 0};
#pragma unroll
        for (int bz = 0; bz < 2; bz++) {
            const int mz = zz - 1 + bz;
            if (mz < 0 || mz >= L) continue;
#pragma unroll
            for (int by = 0; by < 2; by++) {
                const int my = yy - 1 + by;
                if (my < 0 || my >= L) continue;
#pragma unroll
                for (int bx = 0; bx < 2; bx++) {
                    const int mx = xx - 1 + bx;
                    if (mx < 0 || mx >= L) continue;
                    const int m = (mz * L + my) * L + mx;
                    const int j = hex8_corner(1 - bx, 1 - by, 1 - bz);  // the node's corner in micro element m
#pragma unroll
                    for (int ci = 0; ci < 3; ci++) acc[ci] += sY[le2][m][ci * 8 + j];
                }
            }
        }
        if (pos[i] < 0) {
#pragma unroll
            for (int ci = 0; ci < 3; ci++) {
                const ptrdiff_t qd = (ptrdiff_t)gid[i] * 3 + ci;
                if (cs.p) {
                    const T pi = cheb3_direction<T>(cs.first ? T(0) : cs.p[qd], cs.rhs[qd], acc[ci], cs.beta, cs.dinv,
                                                    qd, cs.first);
                    cs.p[qd] = pi;
                    cs.x[qd] = __builtin_fma(-cs.alpha, pi, xk[i][ci]);
                } else {
                    y[qd] = overwrite ? acc[ci] : yk[i][ci] + acc[ci];
                }
            }
        } else {
#pragma unroll
            for (int ci = 0; ci < 3; ci++) part[(ptrdiff_t)pos[i] * 3 + ci] = acc[ci];
        }
    }
}

// d += diagonal (block = 0) or symmetric 3x3 block diagonal [node][6] (block = 1).
template <typename T, int L>
__global__ void sshex8_elasticity_diag_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                              const double *__restrict__ Kall, const int block, T *__restrict__ d) {
    constexpr int N1 = L + 1, NXE = N1 * N1 * N1;
    const ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * NXE) return;
    const ptrdiff_t e = t / NXE;
    const int v = (int)(t - e * NXE);
    const int xx = v % N1, yy = (v / N1) % N1, zz = v / (N1 * N1);
    const double *K = Kall + e * 576;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 8; c++) {
        const int cx = hex8_cx(c), cy = hex8_cy(c), cz = hex8_cz(c);
        const bool ok = (cx ? xx > 0 : xx < L) && (cy ? yy > 0 : yy < L) && (cz ? zz > 0 : zz < L);
        if (!ok) continue;
        for (int k = 0; k < 6; k++) acc[k] += K[(ca[k] * 8 + c) * 24 + cb[k] * 8 + c];
    }
    const ptrdiff_t g = idx[e * NXE + v];
    if (block) {
        for (int k = 0; k < 6; k++) atomic_add(&d[g * 6 + k], (T)acc[k]);
    } else {
        atomic_add(&d[g * 3 + 0], (T)acc[0]);
        atomic_add(&d[g * 3 + 1], (T)acc[3]);
        atomic_add(&d[g * 3 + 2], (T)acc[5]);
    }
}

// ------------------------------------------------------------------ dispatch
#define SFEM_AMD_SSHEX8_LEVELS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16)

bool sshex8_level_supported(int L) {
    switch (L) {
#define CASE(LV) \
    case LV:     \
        return true;
        SFEM_AMD_SSHEX8_LEVELS(CASE)
#undef CASE
        default:
            return false;
    }
}

template <typename T>
int skeleton_gather(const SkeletonMap &sk, int bs, const T *part, T *y, hipStream_t s, int overwrite,
                    int node_sorted, const T *cx, const ChebStep<T> &cs) {
    if (sk.n == 0) return SFEM_SUCCESS;
    // node-sorted partials are read in place; element-major ones through ent.  Default: scalar
    // partials node-sorted, vector partials element-major
    if (node_sorted < 0) node_sorted = bs == 1;
    const int32_t *ent = node_sorted ? nullptr : sk.ent;
    const uint8_t *cm = sk.cmask;
    if (bs == 1)
        skeleton_gather_kernel<T, 1>
                <<<grid_for(sk.n), kBlock, 0, s>>>(sk.n, sk.node, sk.off, ent, part, y, overwrite, cm, cx, cs);
    else if (bs == 3)
        skeleton_gather_kernel<T, 3>
                <<<grid_for(sk.n), kBlock, 0, s>>>(sk.n, sk.node, sk.off, ent, part, y, overwrite, cm, cx, cs);
    else {
        set_last_error("skeleton_gather: block size must be 1 or 3");
        return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T, typename G>
int sshex8_laplacian_apply(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                           const float *coeff, const double *Kem, const T *x, T *y, T *part, const SkeletonMap &sk,
                           hipStream_t s, int overwrite, bool atomic_skeleton) {
    if (E == 0) return SFEM_SUCCESS;
    // atomic skeleton sums: accumulating, unconstrained applies only
    const bool atomic = atomic_skeleton && !overwrite && !sk.cmask;
    if (atomic) part = nullptr;
    switch (L) {
#define CASE(LV)                                                                                                  \
    case LV: {                                                                                                    \
        using Sh = LapShape<LV>;                                                                                  \
        const unsigned grid = (unsigned)((E + Sh::EPB - 1) / Sh::EPB);                                            \
        if constexpr (lap_staged_used<T, LV>())                                                                   \
            sshex8_laplacian_apply_staged_kernel<T, G, LV><<<grid, Sh::BLOCK, 0, s>>>(                            \
                E, idx, fff, fstride, coeff, Kem, x, y, part, sk.pos, overwrite);                                 \
        else                                                                                                      \
            sshex8_laplacian_apply_kernel<T, G, LV><<<grid, Sh::BLOCK, 0, s>>>(E, idx, fff, fstride, coeff, Kem, x, y, \
                                                                                 part, sk.pos, overwrite);        \
        break;                                                                                                    \
    }
        SFEM_AMD_SSHEX8_LEVELS(CASE)
#undef CASE
        default:
            set_last_error("sshex8 level not supported: " + std::to_string(L));
            return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    if (atomic) return SFEM_SUCCESS;
    return skeleton_gather<T>(sk, 1, part, y, s, overwrite, -1, x);
}

template <typename T, typename G>
int sshex8_laplacian_diag(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                          const float *coeff, const double *Kem, T *d, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    switch (L) {
#define CASE(LV)                                                                                          \
    case LV: {                                                                                            \
        using Sh = LapShape<LV>;                                                                          \
        const unsigned grid = (unsigned)((E + Sh::EPB - 1) / Sh::EPB);                                    \
        sshex8_laplacian_diag_kernel<T, G, LV><<<grid, Sh::BLOCK, 0, s>>>(E, idx, fff, fstride, coeff, Kem, d); \
        break;                                                                                            \
    }
        SFEM_AMD_SSHEX8_LEVELS(CASE)
#undef CASE
        default:
            set_last_error("sshex8 level not supported: " + std::to_string(L));
            return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename G>
int sshex8_elasticity_matrices(int L, ptrdiff_t E, const G *adj, const G *det, ptrdiff_t gstride, double mu,
                               double lambda, const float *coeff, double *K, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    const ptrdiff_t n = E * 576;
    sshex8_elasticity_matrices_kernel<G>
            <<<(unsigned)((n + 255) / 256), 256, 0, s>>>(E, L, adj, det, gstride, mu, lambda, coeff, K);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

// connectivity of the sub-blocks: macro element e of level L = s * Ls as s^3 elements of level Ls,
// sub-element e * s^3 + (oz * s + oy) * s + ox
__global__ void sshex8_subblock_idx_kernel(const ptrdiff_t E, const int L, const int Ls,
                                           const int32_t *__restrict__ idx, int32_t *__restrict__ out) {
    const int sb = L / Ls, n1 = Ls + 1, nxs = n1 * n1 * n1, nsub = sb * sb * sb;
    const ptrdiff_t n = E * nsub * nxs;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t es = t / nxs;
        const int v = (int)(t - es * nxs);
        const ptrdiff_t e = es / nsub;
        const int o = (int)(es - e * nsub);
        const int ox = o % sb, oy = (o / sb) % sb, oz = o / (sb * sb);
        const int x = v % n1, y = (v / n1) % n1, z = v / (n1 * n1);
        out[t] = idx[e * nxe_of(L) + lidx(L, Ls * ox + x, Ls * oy + y, Ls * oz + z)];
    }
}

int sshex8_subblock_idx(ptrdiff_t E, int L, int Ls, const int32_t *idx, int32_t *out, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    if (Ls <= 0 || L % Ls) {
        set_last_error("sshex8_subblock_idx: the sub-level must divide the level");
        return SFEM_FAILURE;
    }
    const int sb = L / Ls;
    sshex8_subblock_idx_kernel<<<grid_for(E * sb * sb * sb * nxe_of(Ls)), kBlock, 0, s>>>(E, L, Ls, idx, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

bool sshex8_elasticity_cheb3_fusable(int L, bool fp64, bool matrix_cores) {
    if (!matrix_cores) return false;
    return fp64 ? (L == 2 || L == 4 || L == 6 || L == 8) : (L % 2 == 0 && L >= 2 && L <= 12);
}

template <typename T>
int sshex8_elasticity_apply(int L, ptrdiff_t E, const int32_t *idx, const T *K, const T *x, T *y, T *part,
                            const SkeletonMap &sk, hipStream_t s, int overwrite, int matrix_cores, int kshift,
                            const uint32_t *Apk, const ChebStep<T> &cs) {
    if (cs.p && (!overwrite || !sshex8_elasticity_cheb3_fusable(L, std::is_same<T, double>::value, matrix_cores))) {
        set_last_error("sshex8_elasticity_apply: the Chebyshev epilogue needs an overwriting matrix-core apply");
        return SFEM_FAILURE;
    }
    if (E == 0) return SFEM_SUCCESS;
    if (kshift && !(matrix_cores && ((std::is_same<T, double>::value && (L == 2 || L == 4 || L == 6 || L == 8)) ||
                                     (std::is_same<T, float>::value && L % 2 == 0 && L <= 12)))) {
        set_last_error("sshex8_elasticity_apply: shared matrices need the matrix-core kernels");
        return SFEM_FAILURE;
    }
    // L = 2 on the all-at-once kernel (37.8 vs 110 us FP64, 24 vs 47 us FP32 for 32^3 macro elements);
    // at L = 4 its 64-entry result table costs more LDS occupancy than the colours save (176 vs
    // 135 us FP64), so L >= 4 keeps the colour kernels
    if (matrix_cores && L == 2) {
        if (!std::is_same<T, double>::value && !Apk) {
            set_last_error("sshex8_elasticity_apply: the FP32 matrix-core kernel needs packed matrices");
            return SFEM_FAILURE;
        }
        sshex8_elasticity_apply_small_kernel<T, 2><<<(unsigned)((E + 3) / 4), 256, 0, s>>>(
                E, idx, K, (const el_u4 *)Apk, x, y, part, overwrite, kshift, sk.pos, cs);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        return skeleton_gather<T>(sk, 3, part, y, s, overwrite, 1, x, cs);  // node-sorted partials
    }
    if constexpr (std::is_same<T, double>::value) {
        if (matrix_cores && (L == 2 || L == 4 || L == 6 || L == 8)) {
            switch (L) {
#define MCASE(LV)                                                                                           \
    case LV:                                                                                                \
        sshex8_elasticity_apply_mfma64_kernel<LV>                                                           \
                <<<(unsigned)((E + ElMfma<LV>::EPB - 1) / ElMfma<LV>::EPB), 256, 0, s>>>(E, idx, K, x, y, part, \
                                                                                         overwrite, kshift, sk.pos, cs); \
        break;
                MCASE(2) MCASE(4) MCASE(6) MCASE(8)
#undef MCASE
            }
            SFEM_AMD_HIP_CHECK(hipGetLastError());
            return skeleton_gather<T>(sk, 3, part, y, s, overwrite, 1, x, cs);  // node-sorted partials
        }
    }
    if constexpr (std::is_same<T, float>::value) {
        if (matrix_cores && (L == 2 || L == 4 || L == 6 || L == 8 || L == 10 || L == 12)) {
            if (!Apk) {
                set_last_error("sshex8_elasticity_apply: the FP32 matrix-core kernel needs packed matrices");
                return SFEM_FAILURE;
            }
            switch (L) {
#define MCASE(LV)                                                                                               \
    case LV:                                                                                                    \
        sshex8_elasticity_apply_mfma_kernel<LV><<<(unsigned)((E + ElMfma<LV>::EPB - 1) / ElMfma<LV>::EPB), 256, 0, \
                                                  s>>>(E, idx, (const el_u4 *)Apk, x, y, part, overwrite, kshift, \
                                                       sk.pos, cs);                                             \
        break;
                MCASE(2) MCASE(4) MCASE(6) MCASE(8) MCASE(10) MCASE(12)
#undef MCASE
            }
            SFEM_AMD_HIP_CHECK(hipGetLastError());
            return skeleton_gather<T>(sk, 3, part, y, s, overwrite, 1, x, cs);  // node-sorted partials
        }
    }
    switch (L) {
#define CASE(LV)                                                                                        \
    case LV: {                                                                                          \
        using Sh = ElShape<LV>;                                                                         \
        const unsigned grid = (unsigned)((E + Sh::EPB - 1) / Sh::EPB);                                  \
        sshex8_elasticity_apply_kernel<T, LV><<<grid, Sh::BLOCK, 0, s>>>(E, idx, K, x, y, part, overwrite); \
        break;                                                                                          \
    }
        SFEM_AMD_SSHEX8_LEVELS(CASE)
#undef CASE
        default:
            set_last_error("sshex8 level not supported: " + std::to_string(L));
            return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return skeleton_gather<T>(sk, 3, part, y, s, overwrite, -1, x);
}

// ------------------------------------------------------------------------------------------
// Stencil form with the coefficients in VGPRs.  Every 16-lane row of a wave holds one stencil
// table: lane k keeps coefficients k, k + 16, ... (21 doubles), and each FMA takes its coefficient
// from lane n % 16 of the row through the DPP row broadcast (v_fmac_f64 with row_newbcast), so the
// 243 FMAs of a node need no scalar loads and no per-lane masks.  A row's columns share one kind
// of (x, y) position, whose micro elements (xy quadrants, bit qy*2 + qx) define its table: interior
// columns (all four) and the four faces y = 0, x = L, y = L, x = 0 (two quadrants each, L - 1
// columns).  The four corner columns (one quadrant, one column each) would take a row and a
// table apiece for one node per plane; instead the sweep saves each corner's 2x2 block of columns
// and, after it, the block's threads apply the element matrix to the corner columns' micro
// elements (L per column) from LDS.  Register table blocks: S_m, S_0, S_p; on the z = 0 and
// z = L planes the register slots holding S_0 are reloaded with the in-plane block of the elements
// above (S_0lo) and below (S_0up) the node, and restored after z = 0, so every plane runs the same
// 243 FMAs.  Arithmetic is FP64 for both vector precisions (FP32 vectors
// are converted on load and store: on this chip FP32 FMA without packing runs at the FP64 rate).
template <int K>
__device__ __forceinline__ void fmac_row_bcast(double &acc, double coef, double v) {
    asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" : "+v"(acc) : "v"(coef), "v"(v), "i"(K));
}

template <int L>
struct StDpp {
    static constexpr int N1 = L + 1, PLANE = N1 * N1, H = N1 + 2, P2 = H * H;  // halo-padded plane
    static constexpr int NI = (L - 1) * (L - 1);
    static constexpr int RI = (NI + 15) / 16;  // rows of interior columns
    static constexpr int ROWS = RI + 4;        // + the four face rows
    static constexpr int BLOCK = 64 * ((ROWS + 3) / 4);
    static constexpr int NC = 3 * 81;           // coefficients per kind (S_m, S_0, S_p)
    static constexpr int NT = (3 * 81 + 15) / 16;  // register table per lane (S_m, S_0, S_p)
    static constexpr int NTYPES = 5;  // row kinds: interior, the four faces
};

// quadrant mask of column kind t: 0 interior, 1-4 faces y = 0, x = L, y = L, x = 0, 5-8 corners
// (0, 0), (L, 0), (L, L), (0, L) (the stencil kernel uses kinds 0-4)
__host__ __device__ constexpr inline int st_kind_qmask(int t) {
    constexpr int q[9] = {0b1111, 0b1100, 0b0101, 0b0011, 0b1010, 0b1000, 0b0100, 0b0001, 0b0010};
    return q[t];
}

// The FMAs of neighbour NB (values v of its three components) into the outputs at z = p + 1
// (block S_m), p (block S_0) and p - 1 (block S_p), output component fastest
// within each input component, so consecutive FMAs update different accumulators.
template <int NB, int NT>
__device__ __forceinline__ void st_dpp_nbr(double (&ah)[3], double (&am)[3], double (&al)[3], const double (&tab)[NT],
                                           const double v[3]) {
    fmac_row_bcast<(0 + NB * 9 + 0) % 16>(ah[0], tab[(0 + NB * 9 + 0) / 16], v[0]);
    fmac_row_bcast<(81 + NB * 9 + 0) % 16>(am[0], tab[(81 + NB * 9 + 0) / 16], v[0]);
    fmac_row_bcast<(162 + NB * 9 + 0) % 16>(al[0], tab[(162 + NB * 9 + 0) / 16], v[0]);
    fmac_row_bcast<(0 + NB * 9 + 3) % 16>(ah[1], tab[(0 + NB * 9 + 3) / 16], v[0]);
    fmac_row_bcast<(81 + NB * 9 + 3) % 16>(am[1], tab[(81 + NB * 9 + 3) / 16], v[0]);
    fmac_row_bcast<(162 + NB * 9 + 3) % 16>(al[1], tab[(162 + NB * 9 + 3) / 16], v[0]);
    fmac_row_bcast<(0 + NB * 9 + 6) % 16>(ah[2], tab[(0 + NB * 9 + 6) / 16], v[0]);
    fmac_row_bcast<(81 + NB * 9 + 6) % 16>(am[2], tab[(81 + NB * 9 + 6) / 16], v[0]);
    fmac_row_bcast<(162 + NB * 9 + 6) % 16>(al[2], tab[(162 + NB * 9 + 6) / 16], v[0]);
    fmac_row_bcast<(0 + NB * 9 + 1) % 16>(ah[0], tab[(0 + NB * 9 + 1) / 16], v[1]);
    fmac_row_bcast<(81 + NB * 9 + 1) % 16>(am[0], tab[(81 + NB * 9 + 1) / 16], v[1]);
    fmac_row_bcast<(162 + NB * 9 + 1) % 16>(al[0], tab[(162 + NB * 9 + 1) / 16], v[1]);
    fmac_row_bcast<(0 + NB * 9 + 4) % 16>(ah[1], tab[(0 + NB * 9 + 4) / 16], v[1]);
    fmac_row_bcast<(81 + NB * 9 + 4) % 16>(am[1], tab[(81 + NB * 9 + 4) / 16], v[1]);
    fmac_row_bcast<(162 + NB * 9 + 4) % 16>(al[1], tab[(162 + NB * 9 + 4) / 16], v[1]);
    fmac_row_bcast<(0 + NB * 9 + 7) % 16>(ah[2], tab[(0 + NB * 9 + 7) / 16], v[1]);
    fmac_row_bcast<(81 + NB * 9 + 7) % 16>(am[2], tab[(81 + NB * 9 + 7) / 16], v[1]);
    fmac_row_bcast<(162 + NB * 9 + 7) % 16>(al[2], tab[(162 + NB * 9 + 7) / 16], v[1]);
    fmac_row_bcast<(0 + NB * 9 + 2) % 16>(ah[0], tab[(0 + NB * 9 + 2) / 16], v[2]);
    fmac_row_bcast<(81 + NB * 9 + 2) % 16>(am[0], tab[(81 + NB * 9 + 2) / 16], v[2]);
    fmac_row_bcast<(162 + NB * 9 + 2) % 16>(al[0], tab[(162 + NB * 9 + 2) / 16], v[2]);
    fmac_row_bcast<(0 + NB * 9 + 5) % 16>(ah[1], tab[(0 + NB * 9 + 5) / 16], v[2]);
    fmac_row_bcast<(81 + NB * 9 + 5) % 16>(am[1], tab[(81 + NB * 9 + 5) / 16], v[2]);
    fmac_row_bcast<(162 + NB * 9 + 5) % 16>(al[1], tab[(162 + NB * 9 + 5) / 16], v[2]);
    fmac_row_bcast<(0 + NB * 9 + 8) % 16>(ah[2], tab[(0 + NB * 9 + 8) / 16], v[2]);
    fmac_row_bcast<(81 + NB * 9 + 8) % 16>(am[2], tab[(81 + NB * 9 + 8) / 16], v[2]);
    fmac_row_bcast<(162 + NB * 9 + 8) % 16>(al[2], tab[(162 + NB * 9 + 8) / 16], v[2]);
}

template <typename TIO, int L, int G>
__global__ void __launch_bounds__(StDpp<L>::BLOCK)
    sshex8_elasticity_stencil_dpp_kernel(const ptrdiff_t E, const int32_t *__restrict__ idx,
                                         const double *__restrict__ Kall, const TIO *__restrict__ x,
                                         TIO *__restrict__ y, TIO *__restrict__ part,
                                         const int32_t *__restrict__ sk_pos, const int overwrite) {
    using S = StDpp<L>;
    constexpr int N1 = S::N1, PLANE = S::PLANE, H = S::H, P2 = S::P2, NXE = N1 * PLANE, NSK = sk_count(L),
                  NT = S::NT, NC = S::NC;
    constexpr int ST_SLICE0 = 80, ST_SLICE = 96;  // register-table slots 5..10: entries 80..175
    constexpr int XR = 2 * G;      // x planes in LDS: G being read, G arriving
    constexpr int IR = 3 * G + 1;  // node-id ring: planes z - 1 .. z + 3G - 1 of a step starting at z
    // LDS: the per-kind tables (S_m, S_0, S_p of the five row kinds) are only read once into each
    // lane's registers, so they share their space with the x planes and the id ring that follow them
    constexpr size_t SX_BYTES = sizeof(double) * XR * 3 * (P2 + 1), SG_BYTES = sizeof(int32_t) * IR * S::BLOCK,
                     ST_BYTES = sizeof(double) * S::NTYPES * 243;
    constexpr size_t U_BYTES = ST_BYTES > SX_BYTES + SG_BYTES ? ST_BYTES : SX_BYTES + SG_BYTES;
    __shared__ __align__(16) unsigned char ubuf[U_BYTES];
    // register-table slots 5..10 of each kind (entries 80..175: S_0 and its neighbours in S_m, S_p)
    // with the in-plane block of the z = 0 plane (variant 0: the elements above), of the inner
    // planes (1: both sides) and of the z = L plane (2: below)
    __shared__ double s0v[3][S::NTYPES][ST_SLICE];
    __shared__ double sK[576];
    __shared__ double csave[4][4][3][N1];        // x of each corner's 2x2 block of columns, by plane
    __shared__ double clo[4][L][3], cup[4][L][3];  // corner rows of the micro elements of the corner columns
    __shared__ int32_t cpos[4 * N1];             // skeleton slot positions of the corner nodes
    double(*stab)[243] = reinterpret_cast<double(*)[243]>(ubuf);
    double(*sx)[3][P2 + 1] = reinterpret_cast<double(*)[3][P2 + 1]>(ubuf);           // halo-padded planes
    int32_t(*sg)[S::BLOCK] = reinterpret_cast<int32_t(*)[S::BLOCK]>(ubuf + SX_BYTES);  // ids of planes z, by z % IR

    const ptrdiff_t e = xcd_contiguous_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, r = tid >> 4, k = tid & 15;
    for (int i = tid; i < 576; i += S::BLOCK) sK[i] = Kall[e * 576 + i];
    if (tid < 4 * N1) {
        const int c = tid / N1, z = tid - c * N1;
        cpos[tid] = sk_pos[e * NSK + sk_slot(L, (c == 1 || c == 2) ? L : 0, c >= 2 ? L : 0, z)];
    }

    // this lane's column and its row's kind
    int cx = 1, cy = 1, kind = -1;
    bool active = false, loader = false;
    if (r < S::RI) {
        const int ci = r * 16 + k;
        kind = 0;
        if (ci < S::NI) {
            active = true;
            cx = 1 + ci % (L - 1);
            cy = 1 + ci / (L - 1);
        }
    } else if (r < S::RI + 4) {
        const int f = r - S::RI, t = 1 + min(k, L - 2);
        kind = 1 + f;
        active = k < L - 1;
        // the row's first idle lane stages the x of corner column f ((0,0), (L,0), (L,L), (0,L)): the
        // neighbouring face and interior columns read it from the plane
        loader = k == L - 1;
        if (loader) {
            cx = (f == 1 || f == 2) ? L : 0;
            cy = f >= 2 ? L : 0;
        } else if (f == 0) { cx = t; cy = 0; }
        else if (f == 1) { cx = L; cy = t; }
        else if (f == 2) { cx = t; cy = L; }
        else { cx = 0; cy = t; }
    }
    const bool interior = kind == 0;
    const int32_t *row = idx + e * NXE + cy * N1 + cx;  // node (cx, cy, z) at row[z * PLANE]
    const int me = (cy + 1) * H + cx + 1;               // its cell in the halo-padded plane
    // the corner columns' 2x2 blocks this column belongs to: bit c for corner c ((0,0), (L,0), (L,L),
    // (0,L)), its cell there (dy * 2 + dx) in bits 4 + 2c
    unsigned cmem = 0;
    if (active || loader) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int dx = abs(cx - ((c == 1 || c == 2) ? L : 0)), dy = abs(cy - (c >= 2 ? L : 0));
            if (dx <= 1 && dy <= 1) cmem |= (1u << c) | ((unsigned)(dy * 2 + dx) << (4 + 2 * c));
        }
    }
    // Every lane's column lies in the element (idle lanes sit on a real node too), so the x and id
    // loads need no condition; y is read only where it is updated in place.  Loads and stores share one wait counter on this chip, and a wait while both are
    // pending covers all of them: so each step issues all its loads first, consumes every loaded
    // value (into LDS or the stored data) before it issues its stores, and carries no loaded
    // register into the next step (node ids travel through the thread's own LDS ring).
    TIO x0[G][3];
    int32_t g0[2 * G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        g0[g] = row[(g <= L ? g : L) * PLANE];
        g0[G + g] = row[(G + g <= L ? G + g : L) * PLANE];
#pragma unroll
        for (int c = 0; c < 3; c++) x0[g][c] = x[(ptrdiff_t)g0[g] * 3 + c];
    }
    __syncthreads();  // K staged
    // Table entry (block b: S_m, S_0, S_p; neighbour (dx, dy); row i, column j) of a row kind: the
    // sum, over the node's micro elements that the kind has (z side oz: below for S_m, above for
    // S_p, both for S_0; xy quadrants of st_kind_qmask) and that hold the neighbour, of the element
    // matrix entry coupling them, in the order (oz, qy, qx).  The z = 0 and z = L planes' in-plane
    // blocks: S_0's upper (oz = 1) and lower (oz = 0) sides.  One thread per entry reads the eight
    // micro elements' entries once and sums them for every kind.
    for (int t = tid; t < 243; t += S::BLOCK) {
        const int b = t / 81, q = t - b * 81, nb = q / 9, i = (q % 9) / 3, j = q % 3;
        const int dx = nb % 3 - 1, dy = nb / 3 - 1, dz = b - 1;
        double c[2][4];
        unsigned ok = 0;  // bit oz * 4 + qy * 2 + qx: the neighbour lies in that micro element
#pragma unroll
        for (int oz = 0; oz < 2; oz++)
#pragma unroll
            for (int qy = 0; qy < 2; qy++)
#pragma unroll
                for (int qx = 0; qx < 2; qx++) {
                    const int cz = 1 - oz, cy = 1 - qy, cx = 1 - qx, nz = cz + dz, ny = cy + dy, nx = cx + dx;
                    const bool in = nx >= 0 && nx <= 1 && ny >= 0 && ny <= 1 && nz >= 0 && nz <= 1;
                    c[oz][qy * 2 + qx] =
                            in ? sK[(i * 8 + hex8_corner(cx, cy, cz)) * 24 + j * 8 + hex8_corner(nx, ny, nz)] : 0.0;
                    ok |= (unsigned)in << (oz * 4 + qy * 2 + qx);
                }
#pragma unroll
        for (int kd = 0; kd < S::NTYPES; kd++) {
            const unsigned qmask = (unsigned)st_kind_qmask(kd);
            double v = 0, vside[2] = {0, 0};
#pragma unroll
            for (int oz = 0; oz < 2; oz++) {
                if (!(oz ? b >= 1 : b <= 1)) continue;  // S_m: below only, S_p: above only
#pragma unroll
                for (int qq = 0; qq < 4; qq++)
                    if (((qmask >> qq) & 1u) && ((ok >> (oz * 4 + qq)) & 1u)) {
                        v += c[oz][qq];
                        vside[oz] += c[oz][qq];
                    }
            }
            stab[kd][t] = v;
            if (t >= ST_SLICE0 && t < ST_SLICE0 + ST_SLICE) {
                const bool inplane = b == 1;
                s0v[0][kd][t - ST_SLICE0] = inplane ? vside[1] : v;
                s0v[1][kd][t - ST_SLICE0] = v;
                s0v[2][kd][t - ST_SLICE0] = inplane ? vside[0] : v;
            }
        }
    }
    __syncthreads();  // tables ready
    double tab[NT];
#pragma unroll
    for (int m = 0; m < NT; m++) {
        const int n = m * 16 + k;
        tab[m] = (kind >= 0 && n < 3 * 81) ? stab[kind][n] : 0.0;
    }
    __syncthreads();  // the tables' space becomes the x planes and the id ring
    for (int i = tid; i < XR * 3 * (P2 + 1); i += S::BLOCK) (&sx[0][0][0])[i] = 0;  // the halo stays zero
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; g++) {
        sg[g % IR][tid] = g0[g];
        sg[(G + g) % IR][tid] = g0[G + g];
#pragma unroll
        for (int c = 0; c < 3; c++) sx[g][c][(active || loader) ? me : P2] = (double)x0[g][c];
        if (cmem)
#pragma unroll
            for (int cc = 0; cc < 4; cc++)
                if ((cmem >> cc) & 1u)
#pragma unroll
                    for (int c = 0; c < 3; c++) csave[cc][(cmem >> (4 + 2 * cc)) & 3u][c][g] = (double)x0[g][c];
    }

    double acc_lo[3] = {0, 0, 0}, acc_mid[3] = {0, 0, 0}, acc_hi[3] = {0, 0, 0};
    for (int p0 = 0; p0 <= L; p0 += G) {
        lds_barrier();  // planes p0 .. p0 + G - 1 staged; the other G buffers are free
        // loads of this step: for the nodes at z = p - 1 completed by planes p = p0 .. p0 + G - 1,
        // their old y (interior nodes, accumulating) and skeleton slot positions; the ids of planes
        // p0 + 2G ..; x of planes p0 + G .. (all clamped at z = L)
        TIO yold[G][3], xn[G][3];
        int32_t spos[G], gid2[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int q = p0 + g > 0 ? p0 + g - 1 : 0;
            const int32_t gq = sg[q % IR][tid];
            const int slot = sk_slot(L, cx, cy, q);
            spos[g] = sk_pos[e * NSK + (slot >= 0 ? slot : 0)];
            // only the nodes updated in place read y (skeleton nodes go to the partials)
            const bool read_y = !overwrite && active && interior && q > 0 && q < L;
#pragma unroll
            for (int c = 0; c < 3; c++) yold[g][c] = read_y ? y[(ptrdiff_t)gq * 3 + c] : TIO(0);
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int z2 = p0 + 2 * G + g, z1 = p0 + G + g;
            gid2[g] = row[(z2 <= L ? z2 : L) * PLANE];
            const int32_t g1 = sg[(z1 <= L ? z1 : L) % IR][tid];
#pragma unroll
            for (int c = 0; c < 3; c++) xn[g][c] = x[(ptrdiff_t)g1 * 3 + c];
        }
        double res[G][3];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int p = p0 + g;
            const int buf = p % XR;
            const double *pl0 = &sx[buf][0][0], *pl1 = &sx[buf][1][0], *pl2 = &sx[buf][2][0];
            // all lanes run the FMAs (the table lanes of every row must be active for the broadcast);
            // the outputs of the planes outside [0, L] are never stored.  The z = 0, inner and z = L
            // planes differ in the in-plane block only: its register slots are reloaded on entering
            // them (no table rows are read otherwise)
            if (p == 0 || p == 1 || p == L) {
                const int var = p == 0 ? 0 : (p == L ? 2 : 1);
                const double *sv = s0v[var][kind >= 0 ? kind : 0];
#pragma unroll
                for (int m = ST_SLICE0 / 16; m < (ST_SLICE0 + ST_SLICE) / 16; m++)
                    tab[m] = kind >= 0 && m * 16 + k < 3 * 81 ? sv[m * 16 + k - ST_SLICE0] : 0.0;
            }
            if (p <= L) {
                // the next neighbour's three values are read from LDS before this neighbour's FMAs
                // (volatile asm statements, which the compiler does not move loads across)
#define ST_LD(NB, V)                                                                                    \
    {                                                                                                   \
        constexpr int o = (NB / 3 - 1) * H + (NB % 3 - 1);                                              \
        V[0] = pl0[me + o];                                                                             \
        V[1] = pl1[me + o];                                                                             \
        V[2] = pl2[me + o];                                                                             \
    }
#define ST_NB(NB, V) st_dpp_nbr<NB, NT>(acc_hi, acc_mid, acc_lo, tab, V);
                double va[3], vb[3];
                ST_LD(0, va)
                ST_LD(1, vb) ST_NB(0, va)
                ST_LD(2, va) ST_NB(1, vb)
                ST_LD(3, vb) ST_NB(2, va)
                ST_LD(4, va) ST_NB(3, vb)
                ST_LD(5, vb) ST_NB(4, va)
                ST_LD(6, va) ST_NB(5, vb)
                ST_LD(7, vb) ST_NB(6, va)
                ST_LD(8, va) ST_NB(7, vb)
                ST_NB(8, va)
#undef ST_NB
#undef ST_LD
            }
            // plane p - 1 is complete
#pragma unroll
            for (int c = 0; c < 3; c++) {
                res[g][c] = acc_lo[c];
                acc_lo[c] = acc_mid[c];
                acc_mid[c] = acc_hi[c];
                acc_hi[c] = 0;
            }
        }
        // consume the loads: x of the next G planes into the free buffers (idle lanes: the spare
        // cell) and, for the columns of a corner's 2x2 block, into its save area; the ids into the ring
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int z1 = p0 + G + g;
#pragma unroll
            for (int c = 0; c < 3; c++) sx[z1 % XR][c][(active || loader) ? me : P2] = (double)xn[g][c];
            if (cmem && z1 <= L)
#pragma unroll
                for (int cc = 0; cc < 4; cc++)
                    if ((cmem >> cc) & 1u)
#pragma unroll
                        for (int c = 0; c < 3; c++) csave[cc][(cmem >> (4 + 2 * cc)) & 3u][c][z1] = (double)xn[g][c];
            sg[(p0 + 2 * G + g) % IR][tid] = gid2[g];
        }
        // the stores of the completed planes z = p - 1 (lanes with nothing to store write the
        // 3-entry sink after the partials, which nothing reads: no branch)
#pragma unroll
        for (int g = 0; g < G; g++) {
            // (a step slot p = L + 1 computes nothing and completes z = L)
            const int p = p0 + g, q = p - 1;
            const bool ok = active && p >= 1 && p <= L + 1;
            const bool to_y = interior && q > 0 && q < L;
            const int32_t gq = sg[(q >= 0 ? q : 0) % IR][tid];
            TIO *dst = ok ? (to_y ? y + (ptrdiff_t)gq * 3 : part + (ptrdiff_t)spos[g] * 3) : part + E * NSK * 3;
#pragma unroll
            for (int c = 0; c < 3; c++) dst[c] = (TIO)((to_y ? (double)yold[g][c] : 0.0) + res[g][c]);
        }
    }
    // z = L (a skeleton plane) when the last step ended with it (L + 1 a multiple of G); otherwise the
    // step slot p = L + 1 stored it
    if ((L + 1) % G == 0 && active) {
        const ptrdiff_t ps = (ptrdiff_t)sk_pos[e * NSK + sk_slot(L, cx, cy, L)] * 3;
#pragma unroll
        for (int c = 0; c < 3; c++) part[ps + c] = (TIO)acc_lo[c];
    }
    // the corner columns, in element form from the saved blocks: thread (c, m) applies the rows of
    // the two corner nodes of micro element m of corner column c (its lower and upper corner) to
    // the micro element's 24 values; each corner node's two contributions are summed after
    lds_barrier();  // every wave's saved blocks
    // one row (c, m, lz, i) per thread: 4 corners x L micro elements x 2 corner nodes x 3 components
    for (int t = tid; t < 24 * L; t += S::BLOCK) {
        const int c = t / (6 * L), rem = t - c * 6 * L, m = rem / 6, lz = (rem / 3) & 1, i = rem % 3;
        const int ccx = (c == 1 || c == 2) ? L : 0, ccy = c >= 2 ? L : 0;
        const int mx = ccx ? L - 1 : 0, my = ccy ? L - 1 : 0, lx = ccx - mx, ly = ccy - my;
        const double *kr = sK + (i * 8 + hex8_corner(lx, ly, lz)) * 24;
        double a = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int cell = abs(my + hex8_cy(q) - ccy) * 2 + abs(mx + hex8_cx(q) - ccx), z = m + hex8_cz(q);
#pragma unroll
            for (int j = 0; j < 3; j++) a = fma(kr[j * 8 + q], csave[c][cell][j][z], a);
        }
        (lz ? cup : clo)[c][m][i] = a;
    }
    lds_barrier();
    for (int t = tid; t < 4 * N1; t += S::BLOCK) {
        const int c = t / N1, z = t - c * N1;
        const ptrdiff_t ps = (ptrdiff_t)cpos[t] * 3;
#pragma unroll
        for (int i = 0; i < 3; i++)
            part[ps + i] = (TIO)((z < L ? clo[c][z][i] : 0.0) + (z > 0 ? cup[c][z - 1][i] : 0.0));
    }
}

bool sshex8_elasticity_stencil_supported(int L) {
    return L == 2 || L == 3 || L == 4 || L == 5 || L == 6 || L == 8 || L == 10 || L == 12 || L == 16;
}

template <typename T>
int sshex8_elasticity_apply_stencil(int L, ptrdiff_t E, const int32_t *idx, const double *K64, const T *x, T *y,
                                    T *part, const SkeletonMap &sk, hipStream_t s, int overwrite) {
    if (E == 0) return SFEM_SUCCESS;
    {
        if (!K64 || !part || !sk.pos) {
            set_last_error("sshex8_elasticity_apply_stencil: matrices, partials and slot positions required");
            return SFEM_FAILURE;
        }
        switch (L) {
#define DCASE(LV)                                                                                              \
    case LV:                                                                                                   \
        sshex8_elasticity_stencil_dpp_kernel<T, LV, 1><<<(unsigned)E, StDpp<LV>::BLOCK, 0, s>>>(E, idx, K64, x, y, part, \
                                                                                            sk.pos, overwrite); \
        break;
            DCASE(2) DCASE(3) DCASE(4) DCASE(5) DCASE(6) DCASE(8) DCASE(10) DCASE(12) DCASE(16)
#undef DCASE
            default:
                set_last_error("sshex8 elasticity stencil form: level not supported: " + std::to_string(L));
                return SFEM_FAILURE;
        }
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        return skeleton_gather<T>(sk, 3, part, y, s, overwrite, 1, x);
    }
}

template <typename T>
int sshex8_elasticity_diag(int L, ptrdiff_t E, const int32_t *idx, const double *K, int block, T *d, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    const ptrdiff_t n = E * nxe_of(L);
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (L) {
#define CASE(LV)                                                                          \
    case LV:                                                                              \
        sshex8_elasticity_diag_kernel<T, LV><<<grid, 256, 0, s>>>(E, idx, K, block, d);   \
        break;
        SFEM_AMD_SSHEX8_LEVELS(CASE)
#undef CASE
        default:
            set_last_error("sshex8 level not supported: " + std::to_string(L));
            return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template int skeleton_gather<double>(const SkeletonMap &, int, const double *, double *, hipStream_t, int, int,
                                     const double *, const ChebStep<double> &);
template int skeleton_gather<float>(const SkeletonMap &, int, const float *, float *, hipStream_t, int, int,
                                    const float *, const ChebStep<float> &);
template int sshex8_laplacian_apply<double, float>(int, ptrdiff_t, const int32_t *, const float *, ptrdiff_t,
                                                   const float *, const double *, const double *, double *, double *,
                                                   const SkeletonMap &, hipStream_t, int, bool);
template int sshex8_laplacian_apply<float, uint16_t>(int, ptrdiff_t, const int32_t *, const uint16_t *, ptrdiff_t,
                                                     const float *, const double *, const float *, float *, float *,
                                                     const SkeletonMap &, hipStream_t, int, bool);
template int sshex8_laplacian_diag<double, float>(int, ptrdiff_t, const int32_t *, const float *, ptrdiff_t,
                                                  const float *, const double *, double *, hipStream_t);
template int sshex8_laplacian_diag<float, uint16_t>(int, ptrdiff_t, const int32_t *, const uint16_t *, ptrdiff_t,
                                                    const float *, const double *, float *, hipStream_t);
template int sshex8_elasticity_matrices<float>(int, ptrdiff_t, const float *, const float *, ptrdiff_t, double,
                                               double, const float *, double *, hipStream_t);
template int sshex8_elasticity_matrices<uint16_t>(int, ptrdiff_t, const uint16_t *, const uint16_t *, ptrdiff_t,
                                                  double, double, const float *, double *, hipStream_t);
template int sshex8_elasticity_apply<double>(int, ptrdiff_t, const int32_t *, const double *, const double *,
                                             double *, double *, const SkeletonMap &, hipStream_t, int, int, int,
                                             const uint32_t *, const ChebStep<double> &);
template int sshex8_elasticity_apply<float>(int, ptrdiff_t, const int32_t *, const float *, const float *, float *,
                                            float *, const SkeletonMap &, hipStream_t, int, int, int,
                                            const uint32_t *, const ChebStep<float> &);
template int sshex8_elasticity_apply_stencil<double>(int, ptrdiff_t, const int32_t *, const double *, const double *,
                                                     double *, double *, const SkeletonMap &, hipStream_t, int);
template int sshex8_elasticity_apply_stencil<float>(int, ptrdiff_t, const int32_t *, const double *, const float *,
                                                    float *, float *, const SkeletonMap &, hipStream_t, int);
template int sshex8_elasticity_diag<double>(int, ptrdiff_t, const int32_t *, const double *, int, double *,
                                            hipStream_t);
template int sshex8_elasticity_diag<float>(int, ptrdiff_t, const int32_t *, const double *, int, float *,
                                           hipStream_t);

}  // namespace sfem_amd
