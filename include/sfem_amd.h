/*
 * sfem_amd.h -- C ABI of the MI355X (gfx950) matrix-free FEM + multigrid library.
 *
 * Everything here is extern "C", takes plain pointers and sizes and returns
 * SFEM_SUCCESS (0) or SFEM_FAILURE (1) as the reference does (base/sfem_base.hpp:12-13).
 * Host arrays use the reference's layouts: connectivity SoA int32 [nxe][stride]
 * (the reference's `idx_t **elements` with elements[v] = &elements[v*stride]),
 * FFF float32 AoS [E][6], adjugate float32 AoS [E][9], determinant float32 [E].
 * Vectors passed to apply-like calls are DEVICE pointers; element-type vectors of
 * dofs are interleaved [node*block_size + component] like the reference.
 *
 * Handles:
 *   sfem_amd_mesh  -- connectivity resident in HBM in the device layout
 *                     (element-major for sshex8, SoA for tet4/hex8);
 *   sfem_amd_op    -- anything that maps x -> y: an FE operator, a transfer, a
 *                     solver or multigrid used as an operator
 *                     (the reference's sfem::Operator<T>, algebra/sfem_Operator.hpp:12-33).
 * Operator apply ACCUMULATES, y += A x (frontend/sfem_Function.cpp:481-496); solvers
 * use x as initial guess and output (algebra/sfem_cg.hpp:207-213).
 * `stream` is a hipStream_t passed as void*; NULL is the default stream.  Calls do
 * not synchronise except where they return a host scalar (dot, norm, solve).
 * Objects are not re-entrant: calls that overlap on one handle race.  A mesh must outlive
 * the operators and transfers created on it; operators built from other operators (solvers,
 * constrained operators, multigrid) keep shared references to them.
 */
#ifndef SFEM_AMD_H
#define SFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFEM_SUCCESS 0
#define SFEM_FAILURE 1

/* element types: numeric values of the reference's smesh::ElemType for TET4/HEX8;
 * SSHEX8 is HEX8 refined to `level` (frontend/sfem_FunctionSpace.hpp:17-78). */
enum sfem_amd_element_type { SFEM_AMD_TET4 = 4, SFEM_AMD_HEX8 = 8, SFEM_AMD_SSHEX8 = 108 };

/* value types of an operator's x/y: FP64 (the plain method) or FP32 (chosen form 1) */
enum sfem_amd_real_type { SFEM_AMD_FLOAT64 = 0, SFEM_AMD_FLOAT32 = 1 };

typedef struct sfem_amd_mesh sfem_amd_mesh;
typedef struct sfem_amd_op sfem_amd_op;

/* ------------------------------------------------------------------ library */
/* Last error message of this thread ("" if none). */
const char *sfem_amd_last_error(void);

/* Library options (no counterpart in the reference).  "fused_chebyshev" (default 0): 1 makes
 * Chebyshev3 smoothers on sshex8 matrix-core elasticity operators run each step inside the
 * operator's apply kernels (same results, no product vector; slower on MI355X than the separate
 * streaming update, whose accesses are contiguous); 0 runs the separate product and update.
 * "atomic_skeleton" (default 0): 1 makes accumulating (y += A x), unconstrained sshex8 Laplacian
 * applies add the macro-element boundary contributions to y with atomics, one launch and no
 * partials (results differ from the default in the last bits and are not reproducible run to
 * run); 0 keeps the fixed-order two-phase sum.  "elasticity_form" (default 0): the sshex8 linear
 * elasticity apply in stencil form (a 27-point stencil of 3x3 blocks per macro element, FP64
 * arithmetic) or in element form (the shared 24x24 micro-element matrix, on the matrix cores for even
 * levels); 0 picks the stencil form for FP64 operators at levels >= 10, 1 forces the element form,
 * 2 the stencil form (levels 2-6, 8, 10, 12, 16).  Both give the reference's operator; the
 * summation order differs.  "hex8_patches" (default 1): hex8 elasticity applies on meshes of at least
 * 2^15 elements run on element patches (512 elements grown over shared nodes, accumulated in LDS,
 * patch-boundary nodes summed in a fixed order); 0 runs one element per lane with atomics.
 * "mixed_cg" (default 1): flexible CGs created afterwards around a bridged FP32 preconditioner keep
 * z in FP32 (see sfem_amd_cg_create; the same iterates); 0 runs them through the bridge's casts. */
int sfem_amd_set_option(const char *name, int value);
/* Number of visible GPUs; 0 when no HIP device is present. */
int sfem_amd_device_count(void);
int sfem_amd_device_synchronize(void);

/* ------------------------------------------------------------------ device memory */
void *sfem_amd_malloc(size_t bytes);
int sfem_amd_free(void *ptr);
int sfem_amd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sfem_amd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sfem_amd_memset(void *dst, int value, size_t bytes, void *stream);

/* ------------------------------------------------------------------ host mesh setup */
/* python/sfem/mesh/box_mesh.py:479-700: nx,ny,nz node counts, node id yi*nx*nz+xi*nz+zi,
 * hex8 (8 rows) or 6 tets per cube (4 rows). elements [nv][E], points [3][N] float32. */
ptrdiff_t sfem_amd_box_nnodes(int nx, int ny, int nz);
ptrdiff_t sfem_amd_box_nelements(int nx, int ny, int nz, int element_type);
int sfem_amd_box_mesh(int nx, int ny, int nz, double w, double h, double t, int element_type,
                      int32_t *elements, float *points);
/* Hex8 -> sshex8 level L; replaces smesh::to_semistructured(L, mesh, hierarchical, false)
 * (ssmg/tests/sfem_SSGMGTest.cpp:77-88) with the numbering of the in-tree text
 * (mesh/sfem_hex8_mesh_graph.cpp:448-735 sshex8_generate_elements, :994-1060
 * sshex8_hierarchical_renumbering).  Raw ids: corners keep the macro ids, then edge-interior nodes
 * (edges in CRS (i<j) order, running from the smaller corner id), side-interior nodes (sides in
 * first-processed (element, side) order, walked from the side's smallest corner id), volume nodes.
 * hierarchical != 0: corners keep their ids, the other nodes are renumbered by first encounter over
 * (level 1 < ... < L/2 < L, element, z, y, x), so every derefined level is a prefix of the
 * numbering.  Levels: 1, then L, L/2, ... while even (8: 1 2 4 8; 12: 1 6 12); for odd L > 1,
 * where the reference's renumbering aborts, 1 and L.  With elements == NULL returns the node
 * count only.  sfem_amd_sshex8_refine = the hierarchical numbering. */
ptrdiff_t sfem_amd_sshex8_refine_ex(int level, int hierarchical, ptrdiff_t nelements, ptrdiff_t nnodes,
                                    const int32_t *macro_elements, const float *macro_points,
                                    int32_t *elements, float *points);
ptrdiff_t sfem_amd_sshex8_refine(int level, ptrdiff_t nelements, ptrdiff_t nnodes,
                                 const int32_t *macro_elements, const float *macro_points,
                                 int32_t *elements, float *points);
/* Geometry caches, evaluated in float64 and rounded once to float32:
 * FFF = J^-1 J^-T det (tet4: /6), adjugate and det of J at the element centre.
 * Replaces smesh::FFF / JacobianAdjugateAndDeterminant (frontend/ops/sfem_Laplacian.cpp:88,
 * frontend/ops/sfem_LinearElasticity.cpp:276-278).  For SSHEX8 the macro element's corners
 * are used (level given). */
int sfem_amd_fff(int element_type, int level, ptrdiff_t nelements, ptrdiff_t stride,
                 const int32_t *elements, ptrdiff_t nnodes, const float *points, float *fff);
int sfem_amd_adjugate_det(int element_type, int level, ptrdiff_t nelements, ptrdiff_t stride,
                          const int32_t *elements, ptrdiff_t nnodes, const float *points,
                          float *adjugate, float *det);

/* ------------------------------------------------------------------ device meshes */
/* Upload connectivity (host, reference layout) and keep it in the device layout. */
int sfem_amd_mesh_create(int element_type, int level, ptrdiff_t nelements, ptrdiff_t stride,
                         const int32_t *elements, ptrdiff_t nnodes, sfem_amd_mesh **out);
/* Mesh setup on the device (the reference leaves it to the host smesh library: to_semistructured,
 * ssmg/tests/sfem_SSGMGTest.cpp:77-88; FFF / JacobianAdjugateAndDeterminant,
 * frontend/ops/sfem_Laplacian.cpp:88, frontend/ops/sfem_LinearElasticity.cpp:276-278).
 * sfem_amd_sshex8_refine_device: hex8 macro elements (device, [8][nelements]) -> the sshex8 level-L
 * device mesh with exactly the ids of sfem_amd_sshex8_refine.  _refine_points_device: the refined
 * points (device [3][nnodes of fine]) from the macro points (device [3][nnodes_macro]).
 * sfem_amd_mesh_fff / _adjugate_det: the geometry caches of a device mesh from device points
 * ([3][nnodes of the mesh]) into device arrays [E][6], [E][9] and [E] (as sfem_amd_fff /
 * sfem_amd_adjugate_det).  The operator constructors accept these device arrays directly. */
int sfem_amd_sshex8_refine_device(int level, ptrdiff_t nelements, ptrdiff_t nnodes, const int32_t *macro_elements,
                                  sfem_amd_mesh **out, void *stream);
int sfem_amd_sshex8_refine_points_device(const sfem_amd_mesh *fine, ptrdiff_t nnodes_macro,
                                         const int32_t *macro_elements, const float *macro_points, float *points,
                                         void *stream);
int sfem_amd_mesh_fff(const sfem_amd_mesh *m, const float *points, float *fff, void *stream);
int sfem_amd_mesh_adjugate_det(const sfem_amd_mesh *m, const float *points, float *adjugate, float *det,
                               void *stream);
/* frontend/sfem_API.hpp:1412-1475 sshex8_derefine_element_connectivity, on device.
 * level 1 yields a HEX8 mesh in the standard corner order. */
int sfem_amd_mesh_derefine(const sfem_amd_mesh *fine, int level, sfem_amd_mesh **out);
ptrdiff_t sfem_amd_mesh_nnodes(const sfem_amd_mesh *m);
ptrdiff_t sfem_amd_mesh_nelements(const sfem_amd_mesh *m);
int sfem_amd_mesh_level(const sfem_amd_mesh *m);
int sfem_amd_mesh_element_type(const sfem_amd_mesh *m);
/* Copy the connectivity back in the reference layout ([nxe][nelements], host). */
int sfem_amd_mesh_download(const sfem_amd_mesh *m, int32_t *elements);
void sfem_amd_mesh_destroy(sfem_amd_mesh *m);

/* ------------------------------------------------------------------ FE operators */
/* Laplacian apply: the device counterpart of laplacian_apply_opt (operators/laplacian.cpp:197-236),
 * tet4_laplacian_apply_opt (operators/tet4/tet4_laplacian.cpp:237-285),
 * hex8_laplacian_apply_opt (operators/hex8/hex8_laplacian.cpp:309-345),
 * affine_sshex8_laplacian_stencil_apply_fff (operators/sshex8/sshex8_laplacian.cpp:698-769)
 * and cu_laplacian_apply (operators/cuda/cu_laplacian.hpp).
 * fff: host float32 [E][6]; coeff: host float32 [E] per-element multiplier or NULL.
 * real_type selects FP64 or FP32 vectors; FP32 keeps geometry in BF16 (chosen form 1). */
int sfem_amd_laplacian_create(const sfem_amd_mesh *mesh, const float *fff, const float *coeff,
                              int real_type, sfem_amd_op **out);
/* Stored per-macro sub-element matrices ("ss:em" operators: SemiStructuredEMLaplacian /
 * SemiStructuredEMLinearElasticity, frontend/ops/sfem_SemiStructuredEM*.cpp) applied to every
 * micro element of an sshex8 mesh: the device counterpart of sshex8_stencil_element_matrix_apply
 * (bs 1: matrices [E][64], row-major in hex8 corner order) and ..._apply3 (bs 3: [E][576],
 * index component*8 + node), operators/stencil/sshex8_stencil_element_matrix_apply.cpp:10-160.
 * matrices: host FP64 (copied).  rt: vector precision. */
int sfem_amd_element_matrix_op_create(const sfem_amd_mesh *mesh, const double *matrices, int block_size,
                                      int real_type, sfem_amd_op **out);
/* Linear elasticity (3 dofs/node, interleaved): linear_elasticity_apply_adjugate_{aos,soa}
 * (operators/linear_elasticity.cpp:303-392), affine_hex8_linear_elasticity_apply_adjugate_soa
 * (operators/hex8/hex8_linear_elasticity.cpp:625-760),
 * affine_sshex8_linear_elasticity_apply_macro_adjugate
 * (operators/sshex8/sshex8_linear_elasticity.cpp:275-425), cu_linear_elasticity_apply
 * (operators/cuda/cu_linear_elasticity.hpp).  mu, lambda: Lame parameters; coeff scales both.
 * real_type SFEM_AMD_FLOAT32: FP32 vectors with the adjugate and determinant stored as bf16 words
 * (HEX8: sum-factorised FP32 arithmetic; SSHEX8: per-macro matrices on the matrix cores). */
int sfem_amd_linear_elasticity_create(const sfem_amd_mesh *mesh, const float *adjugate,
                                      const float *det, double mu, double lambda,
                                      const float *coeff, int real_type, sfem_amd_op **out);
/* y += A x (device vectors of the op's real type). */
int sfem_amd_op_apply(sfem_amd_op *op, const void *x, void *y, void *stream);
/* d += diag(A): laplacian_diag (operators/laplacian.cpp:156-195),
 * affine_sshex8_laplacian_diag (operators/sshex8/sshex8_laplacian.cpp:333-435),
 * affine_sshex8_linear_elasticity_diag (sshex8_linear_elasticity.cpp:971-). */
int sfem_amd_op_diag(sfem_amd_op *op, void *d, void *stream);
/* d += sym 3x3 block diagonal [node][6] = (xx,xy,xz,yy,yz,zz) of an elasticity op:
 * affine_sshex8_linear_elasticity_block_diag_sym (sshex8_linear_elasticity.cpp:1102-). */
int sfem_amd_op_block_diag_sym(sfem_amd_op *op, void *d, void *stream);
ptrdiff_t sfem_amd_op_rows(const sfem_amd_op *op);
ptrdiff_t sfem_amd_op_cols(const sfem_amd_op *op);
int sfem_amd_op_real_type(const sfem_amd_op *op);
void sfem_amd_op_destroy(sfem_amd_op *op);

/* Function::apply with constraints (frontend/sfem_Function.cpp:481-496): y += A x, then
 * y[c] = x[c] on constrained dofs c = nodes[i]*block_size + component for each listed
 * component (components: array of n_components ints; NULL = all block components).
 * nodes: host int32. */
int sfem_amd_constrained_op_create(sfem_amd_op *op, ptrdiff_t n_nodes, const int32_t *nodes,
                                   int block_size, int n_components, const int *components,
                                   sfem_amd_op **out);
/* x[c] = value on the constrained dofs of a constrained op (dirichlet.cpp:95-106). */
int sfem_amd_constrained_op_set_value(sfem_amd_op *constrained, double value, void *x, void *stream);
/* diag-style vectors: x[c] = value (f->set_value_to_constrained_dofs). */

/* ------------------------------------------------------------------ transfers */
/* sshex8_prolongate (operators/sshex8/sshex8_interpolate.cpp:259-632) between nested
 * levels of one macro mesh, or sshex8_hierarchical_prolongation (:129-258) when `coarse`
 * is the level-1 HEX8 mesh.  Overwrites the fine vector.  Optional zero constraints on the
 * output (create_gmg_data's apply_zero_constraints, algebra/mg/sfem_GeometricMultigrid.cpp:39-66):
 * constrained: a constrained op on the fine level (or NULL). */
int sfem_amd_prolongation_create(const sfem_amd_mesh *coarse, const sfem_amd_mesh *fine,
                                 int block_size, int real_type, const sfem_amd_op *constrained,
                                 sfem_amd_op **out);
/* sshex8_restrict (:633-767) / sshex8_hierarchical_restriction (:12-128): accumulates into
 * the coarse vector (the caller zeroes it, as Multigrid::cycle does); zeroes the coarse
 * constrained dofs afterwards when `constrained` (a coarse-level constrained op) is given. */
int sfem_amd_restriction_create(const sfem_amd_mesh *fine, const sfem_amd_mesh *coarse,
                                int block_size, int real_type, const sfem_amd_op *constrained,
                                sfem_amd_op **out);

/* y[perm[i]*bs + c] += x[i*bs + c] (a permutation between two numberings of the same nodes,
 * e.g. a box mesh and the hierarchical numbering of a coarser mesh refined onto it); with
 * transpose != 0 the inverse map y[i*bs+c] += x[perm[i]*bs+c].  perm: host int32 [n].
 * constrained (optional): zero those dofs of the output afterwards. */
int sfem_amd_permutation_create(ptrdiff_t n, const int32_t *perm, int bs, int real_type, int transpose,
                                const sfem_amd_op *constrained, sfem_amd_op **out);
/* y += B (A x) through an internal temporary (both of one real type). */
int sfem_amd_chain_create(sfem_amd_op *A, sfem_amd_op *B, sfem_amd_op **out);

/* ------------------------------------------------------------------ BLAS-1 (device) */
/* The reference's BLAS<T> (algebra/sfem_tpl_blas.hpp:11-35) on device vectors of real type
 * `rt` (SFEM_AMD_FLOAT64 / SFEM_AMD_FLOAT32); scalars are passed as double.  Reductions are
 * deterministic (fixed two-stage order, FP64 accumulation) and return host scalars
 * (synchronising the stream). */
int sfem_amd_blas_axpby(int rt, ptrdiff_t n, double alpha, const void *x, double beta, void *y, void *stream);
int sfem_amd_blas_zaxpby(int rt, ptrdiff_t n, double alpha, const void *x, double beta, const void *y, void *z,
                         void *stream);
int sfem_amd_blas_scal(int rt, ptrdiff_t n, double alpha, void *x, void *stream);
int sfem_amd_blas_values(int rt, ptrdiff_t n, double value, void *x, void *stream);
int sfem_amd_blas_copy(int rt, ptrdiff_t n, const void *x, void *y, void *stream);
int sfem_amd_blas_reciprocal(int rt, ptrdiff_t n, double alpha, void *x, void *stream);  /* x = alpha / x */
int sfem_amd_blas_xypaz(int rt, ptrdiff_t n, const void *x, const void *y, double alpha, void *z,
                        void *stream);                                                   /* z = x*y + alpha*z */
double sfem_amd_blas_dot(int rt, ptrdiff_t n, const void *x, const void *y, void *stream);
double sfem_amd_blas_norm2(int rt, ptrdiff_t n, const void *x, void *stream);
/* The dot of sfem_amd_blas_dot written to *result, a DEVICE double, without allocating: no
 * hipMalloc, no pinned host memory, no host synchronisation (so it can run under HIP graph
 * capture or in a loop that must not allocate).  No counterpart in the reference: it exists for
 * callers that cannot allocate.  Same kernels, grid and summation order as sfem_amd_blas_dot, so
 * the two results are the same word.  The partials live in a static per-GPU device array of
 * 1024 doubles (8 KiB, held until the process ends); grids never exceed it, so n is limited only
 * by n >= 0.  Every call on one GPU shares that array: calls that overlap on one GPU (concurrent
 * streams or threads) race. */
int sfem_amd_blas_dot_noalloc(int rt, ptrdiff_t n, const void *x, const void *y, double *result, void *stream);
/* y = (real type rt_to) x */
int sfem_amd_blas_cast(int rt_from, ptrdiff_t n, const void *x, int rt_to, void *y, void *stream);
/* Fused CG step (algebra/sfem_cg.hpp:26-72): x += alpha p, r -= alpha Ap; returns (r,r)
 * when with_rtr != 0 (synchronising), else 0. */
double sfem_amd_cg_update_x_r(int rt, ptrdiff_t n, double alpha, const void *p, const void *Ap, void *x, void *r,
                              int with_rtr, void *stream);
/* The same update, also returning (r, z) of the updated r in *rz (synchronising): the r_new . z_old
 * term of the flexible PCG's beta, read before the preconditioner overwrites z (no copy of z). */
double sfem_amd_cg_update_x_r_dot(int rt, ptrdiff_t n, double alpha, const void *p, const void *Ap, void *x, void *r,
                                  const void *z, double *rz, void *stream);
/* Chebyshev3 iteration_update (algebra/sfem_Chebyshev3.hpp:37-52): p = beta p - rhs + Ax; x -= alpha p */
int sfem_amd_cheb3_update(int rt, ptrdiff_t n, double alpha, double beta, const void *rhs, const void *Ax, void *p,
                          void *x, void *stream);
/* p = beta p + inv_diag (Ax - rhs); x -= alpha p (the Jacobi-scaled Chebyshev3 step). */
int sfem_amd_cheb3_update_diag(int rt, ptrdiff_t n, double alpha, double beta, const void *rhs, const void *Ax,
                               const void *inv_diag, void *p, void *x, void *stream);
/* Chebyshev3 iteration 0 (algebra/sfem_Chebyshev3.hpp:208-217) in one pass: p = p - rhs when p holds
 * A x (have_ax), else p = -rhs (zero initial guess); p *= inv_diag when inv_diag != NULL; x -= alpha p
 * (with a zero initial guess x = -alpha p: x is output only). */
int sfem_amd_cheb3_start(int rt, ptrdiff_t n, double alpha, const void *rhs, const void *inv_diag, int have_ax,
                         void *p, void *x, void *stream);

/* ------------------------------------------------------------------ solvers as operators */
/* sfem::ConjugateGradient<real_t> (algebra/sfem_cg.hpp:80-446): plain CG, or PCG when
 * `precond` != NULL.  mode 0: the reference's PCG with a fixed preconditioner (stops on
 * sqrt(r.z)/sqrt(r0.z0), :430-437); mode 1: flexible PCG (beta from r_new.(z_new - z_old); the
 * reference has none) stopping on the true residual ||b - A x||/||b||; mode 2: fixed-
 * preconditioner PCG stopping on ||r||/||b||.  Applying the returned op solves A x = b
 * (x: initial guess and output).  The op keeps shared references to A and precond, whose
 * overwriting apply it calls for z = P r (a zero fill and the apply, unless the operator writes
 * z directly, as the precision bridge does).  Mode 1 with an FP64 A and a precision bridge over an
 * FP32 preconditioner (or a graph op over one) keeps z in FP32: the CG update also writes the
 * FP32 residual the preconditioner reads, the dots and the direction update read z in FP32, and
 * the bridge's FP32 part is replayed as a graph of its own when precond was a graph op; the
 * iterates are those of the FP64 steps on (double) z, bit for bit (set_option "mixed_cg" 0 runs
 * the FP64 steps through the bridge instead). */
int sfem_amd_cg_create(sfem_amd_op *A, sfem_amd_op *precond, int mode, double rtol,
                       double atol, int max_it, sfem_amd_op **out);
/* sfem::Chebyshev3<real_t> (algebra/sfem_Chebyshev3.hpp:57-303) of degree `max_it`; eig_max
 * from PowerMethod (algebra/sfem_PowerMethod.hpp:18-55) with eigen_solver_max_it iterations
 * (rounded up to odd, like the reference) and tolerance, started from `init_vector` (device;
 * NULL = ones, the reference's init_with_ones).  With a constrained operator pass ones with the
 * constrained dofs zeroed: the identity rows would otherwise pin the estimate at 1. */
int sfem_amd_cheb3_create(sfem_amd_op *A, int max_it, double scale_eig_max, double scale_eig_min,
                          int eigen_solver_max_it, double eigen_solver_tol, int initial_guess_zero,
                          const void *init_vector, sfem_amd_op **out);
double sfem_amd_cheb3_eig_max(const sfem_amd_op *cheb);
/* Chebyshev3 on D^-1 A (extension, not in the reference): as sfem_amd_cheb3_create with the
 * residual scaled by inv_diag (device, n entries, copied) in the recurrence and in the power
 * method.  For coefficients varying by orders of magnitude between elements. */
int sfem_amd_cheb3_diag_create(sfem_amd_op *A, const void *inv_diag, int max_it, double scale_eig_max,
                               double scale_eig_min, int eigen_solver_max_it, double eigen_solver_tol,
                               int initial_guess_zero, const void *init_vector, sfem_amd_op **out);
/* Chebyshev3 (on A, or on D^-1 A with inv_diag non-NULL) whose eig_max comes from at most
 * `lanczos_steps` Lanczos steps, stopping when the estimate changes by at most `tol` relative
 * (extension: the reference's Chebyshev3 uses PowerMethod only, algebra/sfem_Chebyshev3.hpp:165-175,
 * which sfem_amd_cheb3_create keeps).  The scaled process runs in the D inner product, where
 * D^-1 A is self-adjoint.  init_vector as for sfem_amd_cheb3_create (its D-scaled form starts the
 * process). */
int sfem_amd_cheb3_lanczos_create(sfem_amd_op *A, const void *inv_diag, int max_it, double scale_eig_max,
                                  double scale_eig_min, int lanczos_steps, double tol, int initial_guess_zero,
                                  const void *init_vector, sfem_amd_op **out);
/* Iterations the eigenvalue estimate of a Chebyshev3 smoother took (power or Lanczos); -1 for
 * other operators (extension). */
int sfem_amd_cheb3_eigen_iterations(const sfem_amd_op *cheb);
/* StationaryIteration with (scaled) Jacobi: x += omega D^-1 (b - A x), `steps` sweeps
 * (algebra/sfem_Stationary.hpp:50-80, algebra/sfem_ShiftableJacobi.hpp).  diag: device. */
int sfem_amd_jacobi_create(sfem_amd_op *A, const void *diag, double omega, int steps,
                           sfem_amd_op **out);
/* StationaryIteration with block-Jacobi on 3x3 blocks (create_gmg_default_smoothers_and_solver for
 * block size 3, algebra/mg/sfem_GeometricMultigrid.cpp:114-131; ShiftableBlockSymJacobi,
 * algebra/sfem_ShiftableJacobi.hpp): x += omega B^-1 (b - A x), `steps` sweeps.  block_diag_sym:
 * device [n/3][6] in A's real type (sfem_amd_op_block_diag_sym of the unconstrained operator);
 * dof_mask: device uint8 [n], 1 on constrained dofs (their block rows become unit rows), or NULL. */
int sfem_amd_block_jacobi_create(sfem_amd_op *A, const void *block_diag_sym, const uint8_t *dof_mask, double omega,
                                 int steps, sfem_amd_op **out);
/* y += omega * x / diag (create_inverse_diagonal_scaling, used as the coarse-solve preconditioner by
 * drivers/mgsolve.cpp:221-226).  diag: device vector of n entries (copied). */
int sfem_amd_inverse_diagonal_create(int real_type, ptrdiff_t n, const void *diag, double omega,
                                     sfem_amd_op **out);
/* sfem::Multigrid<real_t> (algebra/sfem_Multigrid.hpp:27-347).  Level 0 is the finest;
 * add_level(op, smoother_or_solver, prolongation (NULL on level 0), restriction (NULL on
 * the coarsest)). */
int sfem_amd_mg_create(int real_type, int cycle_type, int max_it, double atol, sfem_amd_op **out);
int sfem_amd_mg_add_level(sfem_amd_op *mg, sfem_amd_op *op, sfem_amd_op *smoother,
                          sfem_amd_op *prolongation, sfem_amd_op *restriction);
/* A view of an operator in the other precision: casts x, applies into a zeroed temporary of
 * the operator's precision, and accumulates the cast result into y.  FP32 op -> FP64 view (the
 * FP32 preconditioner inside FP64 flexible CG, chosen form 1); FP64 op -> FP32 view (the FP64
 * coarse solve inside an FP32 multigrid). */
int sfem_amd_precision_bridge_create(sfem_amd_op *op, sfem_amd_op **out);
/* Solver view in the other precision with in-out x (x is the initial guess and the result):
 * casts b and x, runs the solver/smoother, casts x back.  For an FP64 smoother inside an FP32
 * multigrid. */
int sfem_amd_solver_bridge_create(sfem_amd_op *solver, sfem_amd_op **out);
/* Iterations of the last solve of a solver op; -1 for non-solvers. */
int sfem_amd_op_iterations(const sfem_amd_op *op);
/* Residual history of the last CG solve (relative norms), up to `cap` entries. */
int sfem_amd_cg_history(const sfem_amd_op *op, double *out, int cap);
/* Status of the last solve (SFEM_SUCCESS if converged). */
int sfem_amd_op_last_status(const sfem_amd_op *op);

/* ------------------------------------------------------------------ learned smoother
 * No counterpart in the reference (chosen form 4: a learned sshex8 smoother).  A CNN of `layers`
 * 3x3x3 convolutions without biases (1 -> hidden -> ... -> hidden -> 1 channels, ReLU between)
 * acts on each macro element's (L+1)^3 micro-grid with zero padding:
 *     g = conv_n o relu o ... o relu o conv_1,   f(z) = g(z) - g(-z)
 * Weights are float32 [layer][Cout][Cin][27], offset k = (dz+1)*9 + (dy+1)*3 + (dx+1).
 * weights == NULL: He-normal initialisation from `seed`.  Micro-grid levels 1..8. */
typedef struct sfem_amd_cnn sfem_amd_cnn;
int sfem_amd_cnn_create(int hidden, int layers, const float *weights, unsigned seed, sfem_amd_cnn **out);
/* The same network with `channels` (1 or 3) input and output channels: channels -> hidden -> ... ->
 * channels.  The vector network smooths linear elasticity, one channel per displacement component;
 * its activations and the z / out arrays below are [E][channels][(L+1)^3]; it runs the FP32 path. */
int sfem_amd_cnn_create_channels(int channels, int hidden, int layers, const float *weights, unsigned seed,
                                 sfem_amd_cnn **out);
int sfem_amd_cnn_channels(const sfem_amd_cnn *cnn);
int sfem_amd_cnn_num_params(const sfem_amd_cnn *cnn);
int sfem_amd_cnn_get_weights(const sfem_amd_cnn *cnn, float *weights_host);
int sfem_amd_cnn_set_weights(sfem_amd_cnn *cnn, const float *weights_host);
void sfem_amd_cnn_destroy(sfem_amd_cnn *cnn);
/* out[e][v] = f(z[e][.])[v] for E elements of level L (device, float32 [E][(L+1)^3]).
 * mode 0: FP32 direct convolution; mode 1: fused BF16 MFMA network (hidden 16, 4 layers). */
int sfem_amd_cnn_forward(sfem_amd_cnn *cnn, int level, ptrdiff_t nelements, const float *z, float *out,
                         int mode, void *stream);
/* loss = sum (f(z) - t)^2 / sum t^2 over the batch (device z, t: [E][(L+1)^3]) and its gradient
 * w.r.t. the weights, kept for sfem_amd_cnn_adam_step and copied to grad_host when non-NULL.
 * Synchronises. */
int sfem_amd_cnn_loss_grad(sfem_amd_cnn *cnn, int level, ptrdiff_t nelements, const float *z,
                           const float *target, double *loss, float *grad_host, void *stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
int sfem_amd_cnn_adam_step(sfem_amd_cnn *cnn, double lr, void *stream);
/* out[e][v] = vec[idx[e][v]*bs + comp] * (scale ? scale[idx*bs + comp] : 1) as float32, for an
 * sshex8 mesh (device; vec/scale of real type rt).  comp = -1: every component, as the channels of
 * out [E][bs][(L+1)^3]. */
int sfem_amd_gather_elements(const sfem_amd_mesh *mesh, int real_type, int bs, int comp, const void *vec,
                             const void *scale, float *out, void *stream);
/* Learned smoother on an sshex8 level: `steps` times
 *     r = b - A x;  x += damping * avg_e f(D^-1 r |_e)
 * (for bs > 1: one evaluation of a network with bs channels, or the scalar network on each component;
 * shared nodes average their elements' corrections; constrained dofs
 * untouched when A is a constrained operator).  diag: device diagonal of A (copied).  The op
 * keeps a shared reference to A; mesh and cnn must outlive it.  mode as sfem_amd_cnn_forward. */
int sfem_amd_learned_smoother_create(sfem_amd_op *A, const sfem_amd_mesh *mesh, sfem_amd_cnn *cnn,
                                     const void *diag, int bs, int steps, double damping, int mode,
                                     sfem_amd_op **out);
/* y += M x with a dense row-major FP64 matrix (host, n x n, copied); for the coarsest level of a
 * multigrid hierarchy as a direct solve through a precomputed inverse (no host synchronisation,
 * unlike the reference's coarse CG, drivers/mgsolve.cpp:214-227). */
int sfem_amd_dense_op_create(int real_type, ptrdiff_t n, const double *matrix, sfem_amd_op **out);
/* Wraps `op` so its apply replays a captured HIP graph (no counterpart in the reference): the
 * first apply runs eagerly, the second is captured and launched, later applies with the same
 * (x, y, stream) relaunch the graph and others run eagerly.  For launch-bound sequences without
 * host synchronisation, e.g. a preconditioning V-cycle (max_it 1) with a dense coarse solve; an
 * op whose apply synchronises with the host (a coarse CG) cannot be captured, and its second apply
 * returns the capture error. */
int sfem_amd_graph_op_create(sfem_amd_op *op, sfem_amd_op **out);
/* captured: graphs instantiated (one per initial-guess mode and per accumulating or overwriting
 * apply: the zero-guess overwriting apply a CG makes of its preconditioner records a different
 * sequence), 0 not yet, -1 capture failed (eager); replays: graph launches. */
int sfem_amd_graph_op_stats(const sfem_amd_op *op, int *captured, int *replays);
/* An op that runs `primary` until switched to `fallback` (same size and real type). */
int sfem_amd_switch_create(sfem_amd_op *primary, sfem_amd_op *fallback, sfem_amd_op **out);
int sfem_amd_switch_select(sfem_amd_op *sw, int which);
/* active: 0 primary, 1 fallback; events: switches made by CG fallback hooks. */
int sfem_amd_switch_state(const sfem_amd_op *sw, int *active, int *events);
/* Flexible CG (mode 1) only: after every iteration with ||r_k|| > max_ratio * ||r_{k-1}||, select
 * the fallback of `sw` (once).  Several switches may be registered on one CG. */
int sfem_amd_cg_add_fallback(sfem_amd_op *cg, sfem_amd_op *sw, double max_ratio);

/* ------------------------------------------------------------------ measurement */
/* Times `reps` applies y += A x on `stream` with HIP events (after `warmup` untimed ones).
 * flush_bytes > 0: before every timed apply a kernel overwrites a scratch buffer of that
 * size (outside the timed events) so the apply starts with cold caches.  Returns the
 * total elapsed ms of the timed region (flush_bytes == 0: one event pair around all reps;
 * else the sum of per-apply event pairs) in *total_ms. */
int sfem_amd_op_benchmark(sfem_amd_op *op, const void *x, void *y, int warmup, int reps,
                          size_t flush_bytes, void *stream, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
