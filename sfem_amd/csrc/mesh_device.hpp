// Device-side mesh setup (mesh_device.hip).
#pragma once

#include "common.hpp"
#include "internal.hpp"

namespace sfem_amd {
// hex8 macro elements (device, SoA [8][ne]) -> sshex8 level L connectivity, element-major
// [ne][(L+1)^3] (device), with the ids of the host path sfem_amd_sshex8_refine
int sshex8_refine_device(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *macro_elements,
                         DevArray<int32_t> &idx, ptrdiff_t &ntotal, hipStream_t s);
// refined points [3][nn] (device) from the macro points [3][nn_macro] (device)
int sshex8_refine_points_device(int L, ptrdiff_t ne, ptrdiff_t nn_macro, ptrdiff_t nn, const int32_t *macro_elements,
                                const float *macro_points, const int32_t *idx, float *points, hipStream_t s);
// kind 0: FFF [E][6]; kind 1: adjugate [E][9] and det [E] of a device mesh's (macro) elements
int mesh_geometry_device(int type, int L, ptrdiff_t E, const int32_t *idx, ptrdiff_t nn, const float *points,
                         int kind, float *out, float *det, hipStream_t s);
}  // namespace sfem_amd

namespace sfem_amd {
// per-element words [E][k] (device) -> SoA [k][E] (device); optionally scaled by a per-element
// coefficient (device) and rounded to bf16 (as the host path's to_bf16)
int aos_to_soa_device(const float *aos, ptrdiff_t E, int k, float *soa, hipStream_t s);
int aos_to_bf16_soa_device(const float *aos, const float *coeff, ptrdiff_t E, int k, uint16_t *soa, hipStream_t s);
}  // namespace sfem_amd
