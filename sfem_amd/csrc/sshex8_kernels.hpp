// Host-callable launchers of the sshex8 kernels (sshex8_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {

bool sshex8_level_supported(int L);

// Per-node lists of skeleton slots (slot = e * sk_count(L) + sk_slot(L, x, y, z)): the n shared
// macro-element boundary nodes node[t] (ascending) collect part[ent[off[t] .. off[t+1])].
struct SkeletonMap {
    const int32_t *node{nullptr}, *off{nullptr}, *ent{nullptr};
    ptrdiff_t n{0};
};
// y[node[t]*bs + c] += sum_j part[ent[j]*bs + c]
template <typename T>
int skeleton_gather(const SkeletonMap &sk, int bs, const T *part, T *y, hipStream_t s);

// G = float (FP64 path) or uint16_t holding bf16 (FP32 path).
template <typename T, typename G>
// part: E * sk_count(L) values of scratch (the skeleton contributions)
int sshex8_laplacian_apply(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                           const float *coeff, const T *x, T *y, T *part, const SkeletonMap &sk, hipStream_t s);
template <typename T, typename G>
int sshex8_laplacian_diag(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                          const float *coeff, T *d, hipStream_t s);
template <typename G>
int sshex8_elasticity_matrices(int L, ptrdiff_t E, const G *adj, const G *det, ptrdiff_t gstride, double mu,
                               double lambda, const float *coeff, double *K, hipStream_t s);
template <typename T>
// part: E * sk_count(L) * 3 values of scratch
// K: per-macro 24x24 matrices in the vector precision
int sshex8_elasticity_apply(int L, ptrdiff_t E, const int32_t *idx, const T *K, const T *x, T *y, T *part,
                            const SkeletonMap &sk, hipStream_t s);
template <typename T>
int sshex8_elasticity_diag(int L, ptrdiff_t E, const int32_t *idx, const double *K, int block, T *d, hipStream_t s);

}  // namespace sfem_amd
