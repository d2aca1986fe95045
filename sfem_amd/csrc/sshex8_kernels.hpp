// Host-callable launchers of the sshex8 kernels (sshex8_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace sfem_amd {

bool sshex8_level_supported(int L);

// Skeleton (macro-element boundary) nodes, node[t] ascending, t < n; node[t] collects entries
// off[t] .. off[t+1]-1 of the node-sorted slot list (elements in ascending order).  Scalar
// operators store element e's contribution to its skeleton node (x, y, z) at the node-sorted
// position pos[e * sk_count(L) + sk_slot] of `part`; vector operators store it element-major
// at e * sk_count(L) + sk_slot, and ent[j] is that index for list entry j.
struct SkeletonMap {
    const int32_t *node{nullptr}, *off{nullptr}, *pos{nullptr}, *ent{nullptr};
    ptrdiff_t n{0};
    // optional Dirichlet rows: bit c of cmask[t] marks component c of node[t] as a constrained dof,
    // whose output the gather sets to the operator's input there (y = x) in place of the sum
    const uint8_t *cmask{nullptr};
};
// y[node[t]*bs + c] (+)= sum_{j in [off[t], off[t+1])} part[(bs == 1 ? j : ent[j])*bs + c]
template <typename T>
// node_sorted: partial j of the node-sorted list at part[j * bs] (1) or at part[ent[j] * bs] (0);
// -1: scalar partials node-sorted, vector partials element-major
// cx: the operator's input, read at the constrained rows of sk.cmask (null: those rows are set to
// zero); cs: a
// Chebyshev3 step fed by the (overwritten) product instead of storing y
int skeleton_gather(const SkeletonMap &sk, int bs, const T *part, T *y, hipStream_t s, int overwrite = 0,
                    int node_sorted = -1, const T *cx = nullptr, const ChebStep<T> &cs = ChebStep<T>());

// G = float (FP64 path) or uint16_t holding bf16 (FP32 path).
template <typename T, typename G>
// part: E * sk_count(L) values of scratch (the skeleton contributions)
// Kem != NULL: stored sub-element matrices [E][64] replace the FFF (fff unused)
// atomic_skeleton (accumulating applies without constrained rows): skeleton contributions are
// added to y with atomics (no partials, no gather; summation order not reproducible)
int sshex8_laplacian_apply(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                           const float *coeff, const double *Kem, const T *x, T *y, T *part, const SkeletonMap &sk,
                           hipStream_t s, int overwrite = 0, bool atomic_skeleton = false);
template <typename T, typename G>
int sshex8_laplacian_diag(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                          const float *coeff, const double *Kem, T *d, hipStream_t s);
template <typename G>
int sshex8_elasticity_matrices(int L, ptrdiff_t E, const G *adj, const G *det, ptrdiff_t gstride, double mu,
                               double lambda, const float *coeff, double *K, hipStream_t s);
template <typename T>
// K: per-macro 24x24 matrices in the vector precision; part: E * sk_count(L) * 3 values
// matrix_cores, even L: FP64 (L <= 8) the FP64 MFMA kernel (exact FP64, element-by-element
// summation order); FP32 (L <= 12) the bf16x3 MFMA kernel, ~1e-5 relative per product; otherwise
// the node-wise vector kernel
// kshift > 0 (matrix-core kernels only): elements e share the matrix K[e >> kshift] (sub-blocks)
// Apk: the FP32 matrix-core kernel's operands, sshex8_elasticity_pack_bf16x2 of K (required there)
// cs (matrix-core kernels, overwrite = 1): the product feeds a Chebyshev3 step (x, p updated in
// place; y not written), see sshex8_elasticity_cheb3_fusable
int sshex8_elasticity_apply(int L, ptrdiff_t E, const int32_t *idx, const T *K, const T *x, T *y, T *part,
                            const SkeletonMap &sk, hipStream_t s, int overwrite = 0, int matrix_cores = 0,
                            int kshift = 0, const uint32_t *Apk = nullptr, const ChebStep<T> &cs = ChebStep<T>());
// the level and precision run on the matrix-core kernels (which carry the Chebyshev epilogue)
bool sshex8_elasticity_cheb3_fusable(int L, bool fp64, bool matrix_cores);
// The bf16 hi/lo split of nK float 24 x 24 matrices in the FP32 matrix-core kernel's A-operand
// layout: nK * 1024 words.
int sshex8_elasticity_pack_bf16x2(ptrdiff_t nK, const float *K, uint32_t *Apk, hipStream_t s);
// connectivity [E * (L/Ls)^3][nxe(Ls)] of the macro elements cut into (L/Ls)^3 sub-blocks of level
// Ls (sub-element e * (L/Ls)^3 + (oz * s + oy) * s + ox): the level-L operator as a level-Ls one
int sshex8_subblock_idx(ptrdiff_t E, int L, int Ls, const int32_t *idx, int32_t *out, hipStream_t s);
// Stencil form of the sshex8 elasticity apply (levels 2-6, 8, 10, 12, 16): a 27-point stencil of
// 3x3 blocks per macro element, derived in the kernel from its FP64 24x24 matrix K64 and applied in
// FP64 for either vector precision; part: node-sorted skeleton partials (sk.pos), summed by
// skeleton_gather, followed by 3 spare entries (E * sk_count(L) * 3 + 3 values) that the kernel
// writes and nothing reads.
bool sshex8_elasticity_stencil_supported(int L);
template <typename T>
int sshex8_elasticity_apply_stencil(int L, ptrdiff_t E, const int32_t *idx, const double *K64, const T *x, T *y,
                                    T *part, const SkeletonMap &sk, hipStream_t s, int overwrite = 0);
template <typename T>
int sshex8_elasticity_diag(int L, ptrdiff_t E, const int32_t *idx, const double *K, int block, T *d, hipStream_t s);

}  // namespace sfem_amd
