// Host-callable launchers of transfer_blas_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sfem_amd {

template <typename T>
int sshex8_prolongate(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint8_t *own, int vs, const T *from,
                      T *to, hipStream_t s);
template <typename T>
int sshex8_restrict(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint16_t *count, int vs, const T *from,
                    T *to, hipStream_t s);
// LDS-staged variants (one macro element per block): the prolongation stages the coarse values
// (nxe(Lc) * vs doubles), the restriction the count-scaled fine values (nxe(Lf) * vs doubles) and
// writes coarse skeleton contributions to `part` for skeleton_gather (scalar: node-sorted
// positions cpos; vector: element-major slots), so it adds no atomics.
size_t sshex8_transfer_lds_bytes(int L, int vs);
template <typename T>
// add: to += P from instead of to = P from, leaving the components flagged in keep[node] (bit per
// component, optional) unchanged (a coarse correction added to a fine iterate)
int sshex8_prolongate_lds(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint8_t *own, int vs, const T *from,
                          T *to, hipStream_t s, const uint8_t *keep = nullptr, int add = 0);
template <typename T>
// overwrite: the coarse interior nodes are set, not accumulated (gather the skeleton with overwrite too)
int sshex8_restrict_lds(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint16_t *count, int vs, const T *from,
                        T *to, T *part, const int32_t *cpos, hipStream_t s, int overwrite = 0);
int node_owner_mask(ptrdiff_t E, int nxe, ptrdiff_t nnodes, const int32_t *idx, uint8_t *own, hipStream_t s);
int node_incidence(ptrdiff_t n_entries, ptrdiff_t nnodes, const int32_t *idx, uint16_t *count, hipStream_t s);
int sshex8_derefine(ptrdiff_t E, int Lf, int Lc, const int32_t *fine, int32_t *coarse, hipStream_t s);
int sshex8_to_hex8(ptrdiff_t E, int L, const int32_t *ss, int32_t *h8, hipStream_t s);

template <typename T> int blas_axpby(ptrdiff_t n, T a, const T *x, T b, T *y, hipStream_t s);
template <typename T> int blas_zaxpby(ptrdiff_t n, T a, const T *x, T b, const T *y, T *z, hipStream_t s);
template <typename T> int blas_scal(ptrdiff_t n, T a, T *x, hipStream_t s);
template <typename T> int blas_values(ptrdiff_t n, T a, T *x, hipStream_t s);
template <typename T> int blas_copy(ptrdiff_t n, const T *x, T *y, hipStream_t s);
template <typename T> int blas_reciprocal(ptrdiff_t n, T a, T *x, hipStream_t s);
template <typename T> int blas_xypaz(ptrdiff_t n, const T *x, const T *y, T a, T *z, hipStream_t s);
template <typename S, typename D> int blas_cast(ptrdiff_t n, const S *x, D *y, hipStream_t s);
template <typename T> int cheb3_update(ptrdiff_t n, T alpha, T beta, const T *rhs, const T *Ax, T *p, T *x, hipStream_t s);
// omega * inverse of symmetric 3x3 blocks [n][6] (constrained rows -> unit rows), [n][9]
template <typename T>
int block_sym_inverse(ptrdiff_t n, const T *sym, const uint8_t *mask, double omega, T *inv, hipStream_t s);
template <typename T> int block3_apply(ptrdiff_t n, const T *inv, const T *x, T *y, hipStream_t s);
// y += M x, M dense row-major FP64 n x n
template <typename T> int dense_matvec(ptrdiff_t n, const double *M, const T *x, T *y, hipStream_t s);
template <typename T>
int cheb3_update_diag(ptrdiff_t n, T alpha, T beta, const T *rhs, const T *Ax, const T *dinv, T *p, T *x,
                      hipStream_t s);
// Chebyshev3 iteration 0: p = (have_ax ? p - rhs : -rhs) [* dinv]; x -= alpha p
template <typename T>
int cheb3_start(ptrdiff_t n, T alpha, const T *rhs, const T *dinv, int have_ax, T *p, T *x, hipStream_t s);
// ws: device workspace of reduction_workspace_doubles() doubles
template <typename T> int blas_dot(ptrdiff_t n, const T *x, const T *y, double *ws, double *result_host, hipStream_t s);
// the same dot with its partials in the library's static device scratch: *result (device) = (x, y),
// no allocation and no host synchronisation
template <typename T> int blas_dot_noalloc(ptrdiff_t n, const T *x, const T *y, double *result, hipStream_t s);
template <typename T>
// with z (and rtr_host): also (r, z) of the updated r into *rz_host
int cg_update(ptrdiff_t n, T a, const T *p, const T *Ap, T *x, T *r, double *ws, double *rtr_host, hipStream_t s,
              const T *z = nullptr, double *rz_host = nullptr);
// flexible CG with an FP32 preconditioner: (x, y) for FP64 x and FP32 y; y = a x + b y for FP32 x;
// the CG update that also writes rf = (float) r and, with zf, returns (r, zf) in *rz_host
int blas_dot_mixed(ptrdiff_t n, const double *x, const float *y, double *ws, double *result_host, hipStream_t s);
int blas_axpby_mixed(ptrdiff_t n, double a, const float *x, double b, double *y, hipStream_t s);
int cg_update_mixed(ptrdiff_t n, double a, const double *p, const double *Ap, double *x, double *r, float *rf,
                    const float *zf, double *ws, double *rtr_host, double *rz_host, hipStream_t s);
template <typename T> int dofs_set_value(ptrdiff_t n, const int32_t *dofs, T v, T *x, hipStream_t s);
template <typename T> int dofs_copy(ptrdiff_t n, const int32_t *dofs, const T *src, T *dst, hipStream_t s);
size_t reduction_workspace_doubles();
// y[perm[i]*bs+c] += x[i*bs+c] (transpose: y[i*bs+c] += x[perm[i]*bs+c])
template <typename T>
int permute_add(ptrdiff_t n, int bs, const int32_t *perm, int transpose, const T *x, T *y, hipStream_t s);
// y += (double) x
int cast_accumulate(ptrdiff_t n, const float *x, double *y, hipStream_t s);
// y += (float) x
int cast_accumulate_f32(ptrdiff_t n, const double *x, float *y, hipStream_t s);

}  // namespace sfem_amd
