// Host-callable launchers of the sshex8 kernels (sshex8_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {

bool sshex8_level_supported(int L);

// Skeleton (macro-element boundary) nodes, node[t] ascending, t < n.  The contribution of
// element e to its skeleton node (x, y, z) is stored at position pos[e * sk_count(L) + sk_slot]
// of `part`; the positions of node[t] are off[t] .. off[t+1]-1 (contiguous, in slot order).
struct SkeletonMap {
    const int32_t *node{nullptr}, *off{nullptr}, *pos{nullptr};
    ptrdiff_t n{0};
};
// y[node[t]*bs + c] += sum_{j in [off[t], off[t+1])} part[j*bs + c]
template <typename T>
int skeleton_gather(const SkeletonMap &sk, int bs, const T *part, T *y, hipStream_t s);

// G = float (FP64 path) or uint16_t holding bf16 (FP32 path).
template <typename T, typename G>
// part: E * sk_count(L) values of scratch (the skeleton contributions)
// Kem != NULL: stored sub-element matrices [E][64] replace the FFF (fff unused)
int sshex8_laplacian_apply(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                           const float *coeff, const double *Kem, const T *x, T *y, T *part, const SkeletonMap &sk,
                           hipStream_t s);
template <typename T, typename G>
int sshex8_laplacian_diag(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                          const float *coeff, const double *Kem, T *d, hipStream_t s);
template <typename G>
int sshex8_elasticity_matrices(int L, ptrdiff_t E, const G *adj, const G *det, ptrdiff_t gstride, double mu,
                               double lambda, const float *coeff, double *K, hipStream_t s);
template <typename T>
// part: E * sk_count(L) * 3 values of scratch
// K: per-macro 24x24 matrices in the vector precision; Sst: merged 27-point 3x3-block stencils
// [E][243] (required when sshex8_elasticity_uses_stencils(L))
int sshex8_elasticity_apply(int L, ptrdiff_t E, const int32_t *idx, const T *K, const T *Sst, const T *x, T *y,
                            T *part, const SkeletonMap &sk, hipStream_t s);
bool sshex8_elasticity_uses_stencils(int L);
template <typename T>
int sshex8_elasticity_stencils(ptrdiff_t E, const double *K, T *S, hipStream_t s);
template <typename T>
int sshex8_elasticity_diag(int L, ptrdiff_t E, const int32_t *idx, const double *K, int block, T *d, hipStream_t s);

}  // namespace sfem_amd
