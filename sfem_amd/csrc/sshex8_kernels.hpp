// Host-callable launchers of the sshex8 kernels (sshex8_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {

bool sshex8_level_supported(int L);

// G = float (FP64 path) or uint16_t holding bf16 (FP32 path).
template <typename T, typename G>
int sshex8_laplacian_apply(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                           const float *coeff, const T *x, T *y, hipStream_t s);
template <typename T, typename G>
int sshex8_laplacian_diag(int L, ptrdiff_t E, const int32_t *idx, const G *fff, ptrdiff_t fstride,
                          const float *coeff, T *d, hipStream_t s);
template <typename G>
int sshex8_elasticity_matrices(int L, ptrdiff_t E, const G *adj, const G *det, ptrdiff_t gstride, double mu,
                               double lambda, const float *coeff, double *K, hipStream_t s);
template <typename T>
int sshex8_elasticity_apply(int L, ptrdiff_t E, const int32_t *idx, const double *K, const T *x, T *y,
                            hipStream_t s);
template <typename T>
int sshex8_elasticity_diag(int L, ptrdiff_t E, const int32_t *idx, const double *K, int block, T *d, hipStream_t s);

}  // namespace sfem_amd
