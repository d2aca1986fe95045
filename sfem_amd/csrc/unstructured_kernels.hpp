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
// T = double with float32 geometry words G, or T = float with bf16 words (uint16_t); coeff (nullable)
// scales mu and lambda per element; block: the symmetric 3x3 block diagonal [node][6] instead of the
// diagonal
template <typename T, typename G>
int hex8_elasticity_apply(ptrdiff_t E, const int32_t *idx, const G *adj, const G *det, double mu, double lambda,
                          const float *coeff, const T *x, T *y, hipStream_t s);
// Element patches (built on the host, ElementPatches): elements per patch, and the patch kernels'
// LDS (x and y of every patch node) limit
constexpr int kPatchElements = 512;
constexpr size_t kPatchLdsLimit = 64 * 1024;
inline size_t hex8_patch_lds_bytes(int maxloc, size_t value_bytes) {
    return (size_t)3 * maxloc * (value_bytes + sizeof(double));  // x in the value type, y accumulated in FP64
}
struct PatchView {
    ptrdiff_t npatch{0};
    int maxloc{0};
    const int32_t *count{nullptr}, *node_off{nullptr}, *nodes{nullptr}, *bslot{nullptr};
    const uint16_t *lconn{nullptr};
};
// y += A x over the patches: the patches' own nodes are added to y, the shared ones written to part
// (node-sorted partial slots, 3 values each) for skeleton_gather.  adjp [9][Epad], detp and coeffp
// [Epad] in patch-slot order (patch_gather_rows).
template <typename T, typename G>
int hex8_elasticity_apply_patches(const PatchView &pv, const G *adjp, const G *detp, const float *coeffp, double mu,
                                  double lambda, const T *x, T *y, T *part, hipStream_t s);
// out[k][s] = in[k][perm[s]] for k < rows, s < Epad (0 where perm[s] < 0)
template <typename G>
int patch_gather_rows(int rows, ptrdiff_t E, ptrdiff_t Epad, const int32_t *perm, const G *in, G *out,
                      hipStream_t s);
template <typename T, typename G>
int hex8_elasticity_diag(ptrdiff_t E, const int32_t *idx, const G *adj, const G *det, double mu, double lambda,
                         const float *coeff, int block, T *d, hipStream_t s);
}  // namespace sfem_amd
