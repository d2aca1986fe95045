// Host-callable launchers of the tet4/hex8 kernels (unstructured_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {
int tet4_laplacian_apply(ptrdiff_t E, const int32_t *idx, const float *fff, const double *x, double *y,
                         hipStream_t s);
int tet4_laplacian_diag(ptrdiff_t E, const int32_t *idx, const float *fff, double *d, hipStream_t s);
int hex8_laplacian_apply(ptrdiff_t E, const int32_t *idx, const float *fff, const float *coeff, const double *x,
                         double *y, hipStream_t s);
int hex8_laplacian_diag(ptrdiff_t E, const int32_t *idx, const float *fff, const float *coeff, double *d,
                        hipStream_t s);
int hex8_elasticity_apply(ptrdiff_t E, const int32_t *idx, const float *adj, const float *det, double mu,
                          double lambda, const float *coeff, const double *x, double *y, hipStream_t s);
int hex8_elasticity_diag(ptrdiff_t E, const int32_t *idx, const float *adj, const float *det, double mu,
                         double lambda, const float *coeff, int block, double *d, hipStream_t s);
}  // namespace sfem_amd
