/*
 * sfem_oracle.h -- CPU restatement of SFEM's matrix-free hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (sfem_amd/) never links, calls or falls back to it.
 *
 * Types follow the reference defaults (cmake/SFEMNumbers.cmake:5-14):
 *   idx_t = int32, geom_t = float32, jacobian_t = float32, real_t = float64.
 * Connectivity is SoA: elements[v * stride + e] is local node v of element e,
 * the reference's `idx_t **elements` with elements[v] = &elements[v * stride].
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   box meshes ............ golden fixtures written by importing the reference's
 *                           python/sfem/mesh/box_mesh.py (tests/golden/)
 *   tet4/hex8 geometry .... closed forms (right/scaled tets, parallelepipeds)
 *   element matrices ...... independent Gauss-quadrature assembly in numpy,
 *                           A*1 = 0, rigid-body modes, symmetry, stencil KAT
 *                           (operators/stencil/tests/sfem_StencilTest.cpp:143-216)
 *   sshex8 apply .......... brute-force dense assembly of the refined hex8 mesh
 *   transfers ............. constants prolongate exactly, restriction == P^T,
 *                           Galerkin identity (algebra/sfem_Multigrid.hpp:112-173)
 *   sshex8 numbering ...... known-answer tables derived by hand from the in-tree text
 *                           (mesh/sfem_hex8_mesh_graph.cpp:346-735,994-1060) for one, two and
 *                           permuted elements at L = 2, 3, 4, plus invariants (prefix property,
 *                           counts, coordinates); the live smesh numbering is absent here.
 */
#ifndef SFEM_ORACLE_H
#define SFEM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- meshes ---------------- */

/* Box mesh, python/sfem/mesh/box_mesh.py:479-700. nx,ny,nz are NODE counts per
 * axis (np.linspace(0, w, num=nx)).  Node (xi,yi,zi) has id yi*nx*nz + xi*nz + zi
 * (np.meshgrid 'xy' flattening, leading_dim :66-67).  hex8: 8 SoA rows; tet4: 4 rows,
 * 6 tets per cube (:640-684).  points: [3][nnodes]. */
ptrdiff_t oracle_box_nnodes(int nx, int ny, int nz);
ptrdiff_t oracle_box_nelements(int nx, int ny, int nz, int tet4);
void oracle_box_mesh(int nx, int ny, int nz, double w, double h, double t, int tet4,
                     int32_t *elements, float *points);

/* Number of nodes of the sshex8 refinement of a hex8 mesh (no numbering). */
int  oracle_sshex8_nxe(int L);
/* Refine a hex8 mesh to sshex8 level L: sshex8_generate_elements, then (hierarchical != 0)
 * sshex8_hierarchical_renumbering (mesh/sfem_hex8_mesh_graph.cpp:448-735,994-1060).
 * Out: elements [nxe][nelements], returns node count; points [3][nnodes_out]
 * (elements == NULL: count only).  oracle_sshex8_refine = hierarchical numbering. */
ptrdiff_t oracle_sshex8_refine_ex(int L, int hierarchical, ptrdiff_t nelements, ptrdiff_t nnodes,
                                  const int32_t *macro_elements, const float *macro_points,
                                  int32_t *elements, float *points);
ptrdiff_t oracle_sshex8_refine(int L, ptrdiff_t nelements, ptrdiff_t nnodes,
                               const int32_t *macro_elements, const float *macro_points,
                               int32_t *elements, float *points);
/* sshex8_hierarchical_mesh_levels (mesh/sfem_hex8_mesh_graph.cpp:964-992), coarsest first;
 * returns the count (levels may be NULL).  Odd L > 1: [1, L] (extension). */
int oracle_sshex8_hierarchical_levels(int L, int *levels);
/* frontend/sfem_API.hpp:1412-1475: coarse row lidx(to,x,y,z) = fine row lidx(from, x*s, y*s, z*s). */
void oracle_sshex8_derefine(int from_level, int to_level, ptrdiff_t nelements,
                            const int32_t *fine, int32_t *coarse);
/* Level-1 sshex8 -> standard HEX8 corner order. */
void oracle_sshex8_to_hex8(int L, ptrdiff_t nelements, const int32_t *sselements, int32_t *hex8);
/* Number of elements sharing each node (uint16, like the reference). */
void oracle_node_incidence(int nxe, ptrdiff_t nelements, ptrdiff_t nnodes,
                           const int32_t *elements, uint16_t *count);
/* Max node id + 1 over the given rows. */
ptrdiff_t oracle_max_node(int nxe, ptrdiff_t nelements, const int32_t *elements);

/* ---------------- geometry (float32 words) ---------------- */

/* FFF = J^-1 J^-T det(J)/6 per tet; fff [E][6] AoS.  operators/tet4/tet4_fff.cpp:5-28 */
void oracle_tet4_fff(ptrdiff_t nelements, const int32_t *elements, ptrdiff_t nnodes,
                     const float *points, float *fff);
/* Jacobian at the centre (0.5,0.5,0.5) of the trilinear map, evaluated in float64,
 * rounded on store.  operators/hex8/hex8_inline_cpu.hpp:13-73 (fff),
 * :90- (adjugate, det).  `corners` gives the row of each of the 8 hex8 corners in
 * the connectivity (identity for hex8; sshex8 corner lidx for macro elements). */
void oracle_hex8_fff(ptrdiff_t nelements, const int32_t *elements, const int *corners,
                     ptrdiff_t nnodes, const float *points, float *fff);
void oracle_hex8_adjugate_det(ptrdiff_t nelements, const int32_t *elements, const int *corners,
                              ptrdiff_t nnodes, const float *points, float *adjugate, float *det);

/* ---------------- element matrices ---------------- */
/* Q1 Laplacian 8x8 from FFF (exact integral, constant metric).
 * operators/hex8/hex8_laplacian_inline_cpu.hpp:213-330 */
void oracle_hex8_laplacian_matrix(const double *fff, double *K);
/* P1 Laplacian 4x4.  operators/tet4/tet4_laplacian_inline_cpu.hpp:8-30 */
void oracle_tet4_laplacian_matrix(const double *fff, double *K);
/* Q1 linear elasticity 24x24, index = component*8 + node.
 * operators/hex8/hex8_linear_elasticity_inline_cpu.hpp:207- */
void oracle_hex8_elasticity_matrix(double mu, double lambda, const double *adj, double det, double *K);

/* ---------------- operators: y += A x ---------------- */
/* coeff (nullable): per-element multiplier of the element matrix (variable
 * coefficient extension; NULL reproduces the reference exactly). */
void oracle_tet4_laplacian_apply(ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                 const double *x, double *y);
void oracle_tet4_laplacian_diag(ptrdiff_t nelements, const int32_t *elements, const float *fff, double *d);
void oracle_hex8_laplacian_apply(ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                 const float *coeff, const double *x, double *y);
void oracle_hex8_laplacian_diag(ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                const float *coeff, double *d);
void oracle_sshex8_laplacian_apply(int L, ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                   const float *coeff, const double *x, double *y);
/* "ss:em": stored sub-element matrices applied to every micro element (bs 1: [E][64]; bs 3:
 * [E][576], vectors interleaved).  operators/stencil/sshex8_stencil_element_matrix_apply.cpp:10-160 */
/* the reference's stencil + surface-stencil evaluation of the same operator (the timed CPU path) */
void oracle_sshex8_laplacian_stencil_apply(int L, ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                           const float *coeff, const double *x, double *y);
/* STREAM triad a = b + s c (OpenMP) */
void oracle_stream_triad(ptrdiff_t n, double *a, const double *b, const double *c, double s);
void oracle_sshex8_em_apply(int L, ptrdiff_t nelements, const int32_t *elements, int bs, const double *matrices,
                            const double *x, double *y);
/* operators/sshex8/sshex8_laplacian.cpp:595-634 (from the macro FFF words) */
void oracle_sshex8_laplacian_element_matrices(int L, ptrdiff_t nelements, const float *fff, const float *coeff,
                                              double *out);
void oracle_sshex8_elasticity_element_matrices(int L, ptrdiff_t nelements, const float *adj, const float *det,
                                               double mu, double lambda, const float *coeff, double *out);
void oracle_sshex8_laplacian_diag(int L, ptrdiff_t nelements, const int32_t *elements, const float *fff,
                                  const float *coeff, double *d);
/* elasticity: x,y interleaved [node*3 + c]; adjugate [E][9], det [E]. */
void oracle_hex8_elasticity_apply(ptrdiff_t nelements, const int32_t *elements, const float *adj,
                                  const float *det, double mu, double lambda, const float *coeff,
                                  const double *x, double *y);
void oracle_hex8_elasticity_diag(ptrdiff_t nelements, const int32_t *elements, const float *adj,
                                 const float *det, double mu, double lambda, const float *coeff, double *d);
void oracle_sshex8_elasticity_apply(int L, ptrdiff_t nelements, const int32_t *elements, const float *adj,
                                    const float *det, double mu, double lambda, const float *coeff,
                                    const double *x, double *y);
void oracle_sshex8_elasticity_diag(int L, ptrdiff_t nelements, const int32_t *elements, const float *adj,
                                   const float *det, double mu, double lambda, const float *coeff, double *d);
/* Symmetric 3x3 block diagonal [node][6] = (xx, xy, xz, yy, yz, zz). */
void oracle_sshex8_elasticity_block_diag(int L, ptrdiff_t nelements, const int32_t *elements,
                                         const float *adj, const float *det, double mu, double lambda,
                                         const float *coeff, double *d);

/* ---------------- transfers ---------------- */
/* operators/sshex8/sshex8_interpolate.cpp:259-632 (overwrites `to`). */
void oracle_sshex8_prolongate(ptrdiff_t nelements, int from_level, const int32_t *from_elements,
                              int to_level, const int32_t *to_elements, int vec_size,
                              const double *from, double *to);
/* operators/sshex8/sshex8_interpolate.cpp:633-767 (accumulates into `to`). */
void oracle_sshex8_restrict(ptrdiff_t nelements, int from_level, const int32_t *from_elements,
                            const uint16_t *count, int to_level, const int32_t *to_elements,
                            int vec_size, const double *from, double *to);
/* hex8 (level 1, standard corner order) -> sshex8 level L.  :129-258 (overwrites). */
void oracle_sshex8_hierarchical_prolongation(int L, ptrdiff_t nelements, const int32_t *elements,
                                             int vec_size, const double *from, double *to);
/* :12-128 (accumulates). */
void oracle_sshex8_hierarchical_restriction(int L, ptrdiff_t nelements, const int32_t *elements,
                                            const uint16_t *count, int vec_size,
                                            const double *from, double *to);

/* ---------------- Dirichlet (operators/boundary_conditions/dirichlet.cpp) ---------------- */
void oracle_constraint_nodes_to_value_vec(ptrdiff_t n, const int32_t *nodes, int block_size,
                                          int component, double value, double *x);      /* :95-106 */
void oracle_constraint_nodes_copy_vec(ptrdiff_t n, const int32_t *nodes, int block_size,
                                      int component, const double *src, double *dst);    /* :247-258 */

/* Threads used by the OpenMP loops (for the CPU baseline). */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
