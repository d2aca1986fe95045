/*
 * sfem_oracle.c -- CPU restatement of SFEM's matrix-free hot path (parity checker).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into the product.
 *
 * Each function names the reference file:line whose algorithm it restates.  The
 * loops keep the reference's OpenMP structure (parallel over elements, atomic
 * scatter) so that timing them is a fair stand-in for the reference CPU path.
 * Element matrices are derived here from first principles (tensor products of
 * 1-D Q1 integrals, 2-point Gauss for elasticity); they are pinned against an
 * independent numpy assembly and the reference's stencil known-answer test.
 */
#include "sfem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* =====================================================================
 * Meshes
 * ===================================================================== */

ptrdiff_t oracle_box_nnodes(int nx, int ny, int nz) { return (ptrdiff_t)nx * ny * nz; }

ptrdiff_t oracle_box_nelements(int nx, int ny, int nz, int tet4) {
    ptrdiff_t c = (ptrdiff_t)(nx - 1) * (ny - 1) * (nz - 1);
    return tet4 ? 6 * c : c;
}

/* np.linspace(0, w, num=n, dtype=float32): float64 i*step then cast; last = w. */
static float linspace_f32(int i, int n, double w) {
    if (n == 1) return 0.f;
    if (i == n - 1) return (float)w;
    double step = w / (double)(n - 1);
    return (float)((double)i * step);
}

void oracle_box_mesh(int nx, int ny, int nz, double w, double h, double t, int tet4,
                     int32_t *elements, float *points) {
    /* box_mesh.py:66-67 leading_dim = [nz, nz*nx, 1] */
    const ptrdiff_t ld0 = nz, ld1 = (ptrdiff_t)nz * nx, ld2 = 1;
    const ptrdiff_t nn = oracle_box_nnodes(nx, ny, nz);
    for (int yi = 0; yi < ny; yi++)
        for (int xi = 0; xi < nx; xi++)
            for (int zi = 0; zi < nz; zi++) {
                ptrdiff_t id = yi * ld1 + xi * ld0 + zi * ld2;
                points[0 * nn + id] = linspace_f32(xi, nx, w);
                points[1 * nn + id] = linspace_f32(yi, ny, h);
                points[2 * nn + id] = linspace_f32(zi, nz, t);
            }

    const ptrdiff_t ne = oracle_box_nelements(nx, ny, nz, tet4);
    ptrdiff_t count = 0;
    for (int zi = 0; zi < nz - 1; zi++)
        for (int yi = 0; yi < ny - 1; yi++)
            for (int xi = 0; xi < nx - 1; xi++) {
                const ptrdiff_t x0 = xi * ld0, y0 = yi * ld1, z0 = zi * ld2;
                const ptrdiff_t x1 = (xi + 1) * ld0, y1 = (yi + 1) * ld1, z1 = (zi + 1) * ld2;
                /* box_mesh.py:520-534 corner order */
                const int32_t hx[8] = {(int32_t)(x0 + y0 + z0), (int32_t)(x1 + y0 + z0), (int32_t)(x1 + y1 + z0),
                                       (int32_t)(x0 + y1 + z0), (int32_t)(x0 + y0 + z1), (int32_t)(x1 + y0 + z1),
                                       (int32_t)(x1 + y1 + z1), (int32_t)(x0 + y1 + z1)};
                if (!tet4) {
                    for (int v = 0; v < 8; v++) elements[v * ne + count] = hx[v];
                    count++;
                } else {
                    /* box_mesh.py:640-684: two prisms, three tets each */
                    static const int tets[6][4] = {{0, 1, 3, 7}, {0, 1, 7, 5}, {0, 4, 5, 7},
                                                   {1, 2, 3, 6}, {1, 3, 7, 6}, {1, 5, 6, 7}};
                    for (int k = 0; k < 6; k++) {
                        for (int v = 0; v < 4; v++) elements[v * ne + count] = hx[tets[k][v]];
                        count++;
                    }
                }
            }
}

int oracle_sshex8_nxe(int L) { return (L + 1) * (L + 1) * (L + 1); }

static inline int lidx(int L, int x, int y, int z) { return (z * (L + 1) + y) * (L + 1) + x; }

/* Standard hex8 corner (bits: x=c&1?...) -> structured coordinate. Corner order of
 * box_mesh.py:520-534 / operators/sshex8/sshex8.hpp:53-62. */
static const int hex8_cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int hex8_cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int hex8_cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

/* Local edges (pairs of corners) and faces (4 corners, cyclic) of a hex8. */
static const int hex8_edges[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                                      {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int hex8_faces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                     {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

typedef struct {
    int32_t k[4];
    ptrdiff_t id;
} key4_t;

static int cmp_key4(const void *a, const void *b) {
    const key4_t *p = (const key4_t *)a, *q = (const key4_t *)b;
    for (int i = 0; i < 4; i++) {
        if (p->k[i] != q->k[i]) return p->k[i] < q->k[i] ? -1 : 1;
    }
    return p->id < q->id ? -1 : (p->id > q->id);
}

static void sort4(int32_t *v) {
    for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && v[j - 1] > v[j]; j--) {
            int32_t t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
}

/*
 * sshex8 refinement with hierarchical numbering.  Restates the in-tree
 * historical algorithm (mesh/sfem_hex8_mesh_graph.cpp:346-735, commented out; the
 * live one is smesh::to_semistructured, absent from the reference checkout):
 *   - corners keep their macro ids;
 *   - edge-interior nodes are numbered per unique edge (edges ordered by first
 *     appearance), running from the lower to the higher global corner id;
 *   - face-interior nodes are numbered per unique face (first appearance), with
 *     the origin at the face's min-id corner and u towards its lower-id neighbour;
 *   - volume-interior nodes are base + e*(L-1)^3 + local (x fastest);
 * then the hierarchical renumbering (:994-1080) stably moves nodes that exist at
 * coarser levels (L, L/2, ... down to the odd part, then 1) to the front, so each
 * derefined level is a prefix of the fine numbering.
 */
static int level_rank(int L, int x, int y, int z) {
    /* levels: 1, then the odd part m of L, 2m, 4m, ..., L.  Rank 0 = coarsest. */
    if (x % L == 0 && y % L == 0 && z % L == 0) return 0;
    int m = L;
    while (m % 2 == 0) m /= 2;
    /* level l = m * 2^k has grid step L/l */
    int rank = 1;
    for (int l = m; l <= L; l *= 2, rank++) {
        if (l == 1) continue;
        int s = L / l;
        if (x % s == 0 && y % s == 0 && z % s == 0) return rank;
    }
    return rank;
}

ptrdiff_t oracle_sshex8_refine(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me,
                               const float *mp, int32_t *elements, float *points) {
    const int nxe = oracle_sshex8_nxe(L);
    const int Lm1 = L - 1;

    /* unique edges */
    key4_t *ek = (key4_t *)malloc(sizeof(key4_t) * 12 * (size_t)ne);
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int k = 0; k < 12; k++) {
            int32_t a = me[hex8_edges[k][0] * ne + e], b = me[hex8_edges[k][1] * ne + e];
            key4_t *q = &ek[e * 12 + k];
            q->k[0] = a < b ? a : b;
            q->k[1] = a < b ? b : a;
            q->k[2] = q->k[3] = 0;
            q->id = e * 12 + k;
        }
    qsort(ek, 12 * (size_t)ne, sizeof(key4_t), cmp_key4);
    /* edge id by first appearance: first the minimum occurrence per key */
    ptrdiff_t *occ_edge = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * 12 * (size_t)ne); /* occurrence -> unique slot */
    ptrdiff_t nedges = 0;
    ptrdiff_t *first_occ = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * 12 * (size_t)ne);
    for (ptrdiff_t i = 0; i < 12 * ne; i++) {
        if (i == 0 || ek[i].k[0] != ek[i - 1].k[0] || ek[i].k[1] != ek[i - 1].k[1]) {
            first_occ[nedges] = ek[i].id;
            nedges++;
        }
        occ_edge[ek[i].id] = nedges - 1;
    }
    /* rank unique edges by first occurrence */
    ptrdiff_t *edge_rank = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * (size_t)(nedges + 1));
    {
        key4_t *tmp = (key4_t *)malloc(sizeof(key4_t) * (size_t)(nedges + 1));
        for (ptrdiff_t i = 0; i < nedges; i++) {
            tmp[i].k[2] = tmp[i].k[3] = 0;
            tmp[i].id = i;
            tmp[i].k[0] = (int32_t)(first_occ[i] >> 31);
            tmp[i].k[1] = (int32_t)(first_occ[i] & 0x7fffffff);
        }
        qsort(tmp, nedges, sizeof(key4_t), cmp_key4);
        for (ptrdiff_t i = 0; i < nedges; i++) edge_rank[tmp[i].id] = i;
        free(tmp);
    }

    /* unique faces */
    key4_t *fk = (key4_t *)malloc(sizeof(key4_t) * 6 * (size_t)ne);
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int f = 0; f < 6; f++) {
            key4_t *q = &fk[e * 6 + f];
            for (int i = 0; i < 4; i++) q->k[i] = me[hex8_faces[f][i] * ne + e];
            sort4(q->k);
            q->id = e * 6 + f;
        }
    qsort(fk, 6 * (size_t)ne, sizeof(key4_t), cmp_key4);
    ptrdiff_t *occ_face = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * 6 * (size_t)ne);
    ptrdiff_t nfaces = 0;
    ptrdiff_t *ffirst = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * 6 * (size_t)ne);
    for (ptrdiff_t i = 0; i < 6 * ne; i++) {
        if (i == 0 || memcmp(fk[i].k, fk[i - 1].k, sizeof(int32_t) * 4) != 0) {
            ffirst[nfaces] = fk[i].id;
            nfaces++;
        }
        occ_face[fk[i].id] = nfaces - 1;
    }
    ptrdiff_t *face_rank = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * (size_t)(nfaces + 1));
    {
        key4_t *tmp = (key4_t *)malloc(sizeof(key4_t) * (size_t)(nfaces + 1));
        for (ptrdiff_t i = 0; i < nfaces; i++) {
            tmp[i].k[0] = (int32_t)(ffirst[i] >> 31);
            tmp[i].k[1] = (int32_t)(ffirst[i] & 0x7fffffff);
            tmp[i].k[2] = tmp[i].k[3] = 0;
            tmp[i].id = i;
        }
        qsort(tmp, nfaces, sizeof(key4_t), cmp_key4);
        for (ptrdiff_t i = 0; i < nfaces; i++) face_rank[tmp[i].id] = i;
        free(tmp);
    }

    const ptrdiff_t base_edge = nnodes;
    const ptrdiff_t base_face = base_edge + nedges * Lm1;
    const ptrdiff_t base_vol = base_face + nfaces * (ptrdiff_t)Lm1 * Lm1;
    const ptrdiff_t ntotal = base_vol + ne * (ptrdiff_t)Lm1 * Lm1 * Lm1;

    if (!elements) {
        free(ek); free(occ_edge); free(first_occ); free(edge_rank);
        free(fk); free(occ_face); free(ffirst); free(face_rank);
        return ntotal;
    }

    int32_t *tmp_el = (int32_t *)malloc(sizeof(int32_t) * (size_t)nxe * ne);
    for (ptrdiff_t e = 0; e < ne; e++) {
        int32_t c[8];
        for (int v = 0; v < 8; v++) c[v] = me[v * ne + e];
        for (int z = 0; z <= L; z++)
            for (int y = 0; y <= L; y++)
                for (int x = 0; x <= L; x++) {
                    const int bx = (x == 0 || x == L), by = (y == 0 || y == L), bz = (z == 0 || z == L);
                    const int nb = bx + by + bz;
                    ptrdiff_t id = -1;
                    if (nb == 3) {
                        for (int v = 0; v < 8; v++)
                            if (hex8_cx[v] * L == x && hex8_cy[v] * L == y && hex8_cz[v] * L == z) id = c[v];
                    } else if (nb == 2) {
                        /* edge: find local edge whose two corners bracket this node */
                        for (int k = 0; k < 12; k++) {
                            int a = hex8_edges[k][0], b = hex8_edges[k][1];
                            int ax = hex8_cx[a] * L, ay = hex8_cy[a] * L, az = hex8_cz[a] * L;
                            int bxx = hex8_cx[b] * L, byy = hex8_cy[b] * L;
                            /* node lies on segment a-b? */
                            int t;
                            if (ax != bxx) {
                                if (y != ay || z != az) continue;
                                t = x;
                            } else if (ay != byy) {
                                if (x != ax || z != az) continue;
                                t = y;
                            } else {
                                if (x != ax || y != ay) continue;
                                t = z;
                            }
                            /* t measured from corner a (a has the 0 coordinate on this axis) */
                            int ga = c[a], gb = c[b];
                            int tt = ga < gb ? t : L - t;
                            id = base_edge + edge_rank[occ_edge[e * 12 + k]] * Lm1 + (tt - 1);
                            break;
                        }
                    } else if (nb == 1) {
                        int f;
                        /* face of the node */
                        if (z == 0) f = 0;
                        else if (z == L) f = 1;
                        else if (y == 0) f = 2;
                        else if (y == L) f = 3;
                        else if (x == 0) f = 4;
                        else f = 5;
                        /* origin = min-id corner of the face; u = lower-id neighbour */
                        int fc[4];
                        for (int i = 0; i < 4; i++) fc[i] = hex8_faces[f][i];
                        int am = 0;
                        for (int i = 1; i < 4; i++)
                            if (c[fc[i]] < c[fc[am]]) am = i;
                        int o = fc[am], nu = fc[(am + 1) % 4], nv = fc[(am + 3) % 4];
                        if (c[nu] > c[nv]) {
                            int tmp = nu;
                            nu = nv;
                            nv = tmp;
                        }
                        /* coordinates along u and v, measured from o */
                        int oc[3] = {hex8_cx[o] * L, hex8_cy[o] * L, hex8_cz[o] * L};
                        int uc[3] = {hex8_cx[nu] * L, hex8_cy[nu] * L, hex8_cz[nu] * L};
                        int vc[3] = {hex8_cx[nv] * L, hex8_cy[nv] * L, hex8_cz[nv] * L};
                        int p[3] = {x, y, z};
                        int iu = 0, iv = 0;
                        for (int d = 0; d < 3; d++) {
                            if (uc[d] != oc[d]) iu = (uc[d] > oc[d]) ? p[d] - oc[d] : oc[d] - p[d];
                            if (vc[d] != oc[d]) iv = (vc[d] > oc[d]) ? p[d] - oc[d] : oc[d] - p[d];
                        }
                        id = base_face + face_rank[occ_face[e * 6 + f]] * (ptrdiff_t)Lm1 * Lm1 + (iv - 1) * Lm1 +
                             (iu - 1);
                    } else {
                        id = base_vol + e * (ptrdiff_t)Lm1 * Lm1 * Lm1 + ((z - 1) * Lm1 + (y - 1)) * Lm1 + (x - 1);
                    }
                    tmp_el[lidx(L, x, y, z) * ne + e] = (int32_t)id;
                }
    }

    /* hierarchical renumbering: rank of first level, stable on the id */
    int *rank = (int *)malloc(sizeof(int) * (size_t)ntotal);
    for (ptrdiff_t i = 0; i < ntotal; i++) rank[i] = -1;
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int z = 0; z <= L; z++)
            for (int y = 0; y <= L; y++)
                for (int x = 0; x <= L; x++) rank[tmp_el[lidx(L, x, y, z) * ne + e]] = level_rank(L, x, y, z);
    int maxrank = 0;
    for (ptrdiff_t i = 0; i < ntotal; i++)
        if (rank[i] > maxrank) maxrank = rank[i];
    ptrdiff_t *perm = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * (size_t)ntotal);
    ptrdiff_t next = 0;
    for (int r = 0; r <= maxrank; r++)
        for (ptrdiff_t i = 0; i < ntotal; i++)
            if (rank[i] == r) perm[i] = next++;
    for (ptrdiff_t i = 0; i < (ptrdiff_t)nxe * ne; i++) elements[i] = (int32_t)perm[tmp_el[i]];

    if (points) {
        for (ptrdiff_t e = 0; e < ne; e++) {
            float cx[8], cy[8], cz[8];
            for (int v = 0; v < 8; v++) {
                int32_t n = me[v * ne + e];
                cx[v] = mp[0 * nnodes + n];
                cy[v] = mp[1 * nnodes + n];
                cz[v] = mp[2 * nnodes + n];
            }
            for (int z = 0; z <= L; z++)
                for (int y = 0; y <= L; y++)
                    for (int x = 0; x <= L; x++) {
                        double qx = (double)x / L, qy = (double)y / L, qz = (double)z / L;
                        double f[8] = {(1 - qx) * (1 - qy) * (1 - qz), qx * (1 - qy) * (1 - qz), qx * qy * (1 - qz),
                                       (1 - qx) * qy * (1 - qz),       (1 - qx) * (1 - qy) * qz, qx * (1 - qy) * qz,
                                       qx * qy * qz,                   (1 - qx) * qy * qz};
                        double px = 0, py = 0, pz = 0;
                        for (int v = 0; v < 8; v++) {
                            px += f[v] * cx[v];
                            py += f[v] * cy[v];
                            pz += f[v] * cz[v];
                        }
                        int32_t n = elements[lidx(L, x, y, z) * ne + e];
                        points[0 * ntotal + n] = (float)px;
                        points[1 * ntotal + n] = (float)py;
                        points[2 * ntotal + n] = (float)pz;
                    }
        }
    }

    free(perm); free(rank); free(tmp_el);
    free(ek); free(occ_edge); free(first_occ); free(edge_rank);
    free(fk); free(occ_face); free(ffirst); free(face_rank);
    return ntotal;
}

void oracle_sshex8_derefine(int from_level, int to_level, ptrdiff_t ne, const int32_t *fine, int32_t *coarse) {
    const int s = from_level / to_level;
    for (int z = 0; z <= to_level; z++)
        for (int y = 0; y <= to_level; y++)
            for (int x = 0; x <= to_level; x++) {
                memcpy(&coarse[(size_t)lidx(to_level, x, y, z) * ne], &fine[(size_t)lidx(from_level, x * s, y * s, z * s) * ne],
                       sizeof(int32_t) * ne);
            }
}

void oracle_sshex8_to_hex8(int L, ptrdiff_t ne, const int32_t *ss, int32_t *hex8) {
    for (int v = 0; v < 8; v++) {
        int row = lidx(L, hex8_cx[v] * L, hex8_cy[v] * L, hex8_cz[v] * L);
        memcpy(&hex8[(size_t)v * ne], &ss[(size_t)row * ne], sizeof(int32_t) * ne);
    }
}

void oracle_node_incidence(int nxe, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *el, uint16_t *count) {
    memset(count, 0, sizeof(uint16_t) * nnodes);
    for (int v = 0; v < nxe; v++)
        for (ptrdiff_t e = 0; e < ne; e++) count[el[(size_t)v * ne + e]]++;
}

ptrdiff_t oracle_max_node(int nxe, ptrdiff_t ne, const int32_t *el) {
    ptrdiff_t m = -1;
    for (ptrdiff_t i = 0; i < (ptrdiff_t)nxe * ne; i++)
        if (el[i] > m) m = el[i];
    return m + 1;
}

/* =====================================================================
 * Geometry
 * ===================================================================== */

/* adj(J) with J[i][j] = dx_i/dq_j, row-major; det. */
static void adjugate3(const double *J, double *A, double *det) {
    A[0] = J[4] * J[8] - J[5] * J[7];
    A[1] = J[2] * J[7] - J[1] * J[8];
    A[2] = J[1] * J[5] - J[2] * J[4];
    A[3] = J[5] * J[6] - J[3] * J[8];
    A[4] = J[0] * J[8] - J[2] * J[6];
    A[5] = J[2] * J[3] - J[0] * J[5];
    A[6] = J[3] * J[7] - J[4] * J[6];
    A[7] = J[1] * J[6] - J[0] * J[7];
    A[8] = J[0] * J[4] - J[1] * J[3];
    *det = J[0] * A[0] + J[1] * A[3] + J[2] * A[6];
}

/* FFF_ab = sum_k Jinv[a][k] Jinv[b][k] * det * scale, Jinv = adj/det. */
static void fff_from_jacobian(const double *J, double scale, float *fff) {
    double A[9], det;
    adjugate3(J, A, &det);
    double g[6];
    int k = 0;
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double s = 0;
            for (int c = 0; c < 3; c++) s += A[a * 3 + c] * A[b * 3 + c];
            g[k++] = s / det * scale;
        }
    for (int i = 0; i < 6; i++) fff[i] = (float)g[i];
}

/* operators/tet4/tet4_fff.cpp:5-28 + tet4_inline_cpu.hpp:224-262.  The reference
 * evaluates in float32; here float64 then one rounding (differs by a few ulp). */
void oracle_tet4_fff(ptrdiff_t ne, const int32_t *el, ptrdiff_t nn, const float *p, float *fff) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        int32_t v[4];
        for (int i = 0; i < 4; i++) v[i] = el[(size_t)i * ne + e];
        double J[9];
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < 3; j++) J[d * 3 + j] = (double)p[d * nn + v[j + 1]] - (double)p[d * nn + v[0]];
        fff_from_jacobian(J, 1.0 / 6.0, &fff[e * 6]);
    }
}

/* Trilinear map derivative at the element centre. */
static void hex8_center_jacobian(const double *x, const double *y, const double *z, double *J) {
    const double *c[3] = {x, y, z};
    for (int d = 0; d < 3; d++)
        for (int j = 0; j < 3; j++) J[d * 3 + j] = 0;
    for (int v = 0; v < 8; v++) {
        /* dphi_v/dq at (0.5,0.5,0.5) = (+-1/4, +-1/4, +-1/4) */
        double g[3] = {(hex8_cx[v] ? 0.25 : -0.25), (hex8_cy[v] ? 0.25 : -0.25), (hex8_cz[v] ? 0.25 : -0.25)};
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < 3; j++) J[d * 3 + j] += c[d][v] * g[j];
    }
}

static void gather_corners(ptrdiff_t e, ptrdiff_t ne, const int32_t *el, const int *corners, ptrdiff_t nn,
                           const float *p, double *x, double *y, double *z) {
    for (int v = 0; v < 8; v++) {
        int row = corners ? corners[v] : v;
        int32_t n = el[(size_t)row * ne + e];
        x[v] = p[0 * nn + n];
        y[v] = p[1 * nn + n];
        z[v] = p[2 * nn + n];
    }
}

void oracle_hex8_fff(ptrdiff_t ne, const int32_t *el, const int *corners, ptrdiff_t nn, const float *p, float *fff) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double x[8], y[8], z[8], J[9];
        gather_corners(e, ne, el, corners, nn, p, x, y, z);
        hex8_center_jacobian(x, y, z, J);
        fff_from_jacobian(J, 1.0, &fff[e * 6]);
    }
}

void oracle_hex8_adjugate_det(ptrdiff_t ne, const int32_t *el, const int *corners, ptrdiff_t nn, const float *p,
                              float *adj, float *det) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double x[8], y[8], z[8], J[9], A[9], d;
        gather_corners(e, ne, el, corners, nn, p, x, y, z);
        hex8_center_jacobian(x, y, z, J);
        adjugate3(J, A, &d);
        for (int i = 0; i < 9; i++) adj[e * 9 + i] = (float)A[i];
        det[e] = (float)d;
    }
}

/* =====================================================================
 * Element matrices
 * ===================================================================== */

/* 1-D Q1 integrals on [0,1]: mass M, stiffness S, and C[i][j] = int psi_i' psi_j. */
static const double M1[2][2] = {{1.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 1.0 / 3.0}};
static const double S1[2][2] = {{1.0, -1.0}, {-1.0, 1.0}};
static const double C1[2][2] = {{-0.5, -0.5}, {0.5, 0.5}};

/* int d_a phi_i d_b phi_j over the unit cube */
static double q1_grad_integral(int a, int b, int i, int j) {
    const int bi[3] = {hex8_cx[i], hex8_cy[i], hex8_cz[i]};
    const int bj[3] = {hex8_cx[j], hex8_cy[j], hex8_cz[j]};
    double r = 1;
    for (int d = 0; d < 3; d++) {
        double t;
        if (d == a && d == b) t = S1[bi[d]][bj[d]];
        else if (d == a) t = C1[bi[d]][bj[d]];
        else if (d == b) t = C1[bj[d]][bi[d]];
        else t = M1[bi[d]][bj[d]];
        r *= t;
    }
    return r;
}

void oracle_hex8_laplacian_matrix(const double *fff, double *K) {
    const double G[3][3] = {{fff[0], fff[1], fff[2]}, {fff[1], fff[3], fff[4]}, {fff[2], fff[4], fff[5]}};
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            double s = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) s += G[a][b] * q1_grad_integral(a, b, i, j);
            K[i * 8 + j] = s;
        }
}

/* P1: grad phi = (-1,-1,-1), e1, e2, e3 on the reference tet; FFF carries 1/6. */
void oracle_tet4_laplacian_matrix(const double *fff, double *K) {
    const double G[3][3] = {{fff[0], fff[1], fff[2]}, {fff[1], fff[3], fff[4]}, {fff[2], fff[4], fff[5]}};
    const double g[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) s += g[i][a] * G[a][b] * g[j][b];
            K[i * 4 + j] = s;
        }
}

/* a(u,v) = int lambda div u div v + 2 mu eps(u):eps(v), constant J, 2-point Gauss. */
void oracle_hex8_elasticity_matrix(double mu, double lambda, const double *adj, double det, double *K) {
    const double gp[2] = {0.5 - 0.5 / sqrt(3.0), 0.5 + 0.5 / sqrt(3.0)};
    memset(K, 0, sizeof(double) * 576);
    for (int qz = 0; qz < 2; qz++)
        for (int qy = 0; qy < 2; qy++)
            for (int qx = 0; qx < 2; qx++) {
                const double q[3] = {gp[qx], gp[qy], gp[qz]};
                double g[8][3];
                for (int n = 0; n < 8; n++) {
                    const int b[3] = {hex8_cx[n], hex8_cy[n], hex8_cz[n]};
                    double rg[3];
                    for (int d = 0; d < 3; d++) {
                        double r = 1;
                        for (int k = 0; k < 3; k++) {
                            if (k == d) r *= b[k] ? 1.0 : -1.0;
                            else r *= b[k] ? q[k] : 1 - q[k];
                        }
                        rg[d] = r;
                    }
                    /* physical gradient: d phi/d x_k = sum_j adj[j][k]/det * dphi/dq_j */
                    for (int k = 0; k < 3; k++) {
                        double s = 0;
                        for (int j = 0; j < 3; j++) s += adj[j * 3 + k] * rg[j];
                        g[n][k] = s / det;
                    }
                }
                const double w = det / 8.0;
                for (int ci = 0; ci < 3; ci++)
                    for (int i = 0; i < 8; i++)
                        for (int cj = 0; cj < 3; cj++)
                            for (int j = 0; j < 8; j++) {
                                double dot = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
                                double v = lambda * g[i][ci] * g[j][cj] + mu * (g[i][cj] * g[j][ci]);
                                if (ci == cj) v += mu * dot;
                                K[(ci * 8 + i) * 24 + cj * 8 + j] += w * v;
                            }
            }
}

/* =====================================================================
 * Operators
 * ===================================================================== */

static inline double coef(const float *c, ptrdiff_t e) { return c ? (double)c[e] : 1.0; }

/* operators/tet4/tet4_laplacian.cpp:237-285 */
void oracle_tet4_laplacian_apply(ptrdiff_t ne, const int32_t *el, const float *fff, const double *x, double *y) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double f[6], K[16], u[4];
        int32_t ev[4];
        for (int k = 0; k < 6; k++) f[k] = fff[e * 6 + k];
        oracle_tet4_laplacian_matrix(f, K);
        for (int v = 0; v < 4; v++) {
            ev[v] = el[(size_t)v * ne + e];
            u[v] = x[ev[v]];
        }
        for (int i = 0; i < 4; i++) {
            double s = 0;
            for (int j = 0; j < 4; j++) s += K[i * 4 + j] * u[j];
#pragma omp atomic update
            y[ev[i]] += s;
        }
    }
}

void oracle_tet4_laplacian_diag(ptrdiff_t ne, const int32_t *el, const float *fff, double *d) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double f[6], K[16];
        for (int k = 0; k < 6; k++) f[k] = fff[e * 6 + k];
        oracle_tet4_laplacian_matrix(f, K);
        for (int i = 0; i < 4; i++) {
#pragma omp atomic update
            d[el[(size_t)i * ne + e]] += K[i * 4 + i];
        }
    }
}

/* operators/hex8/hex8_laplacian.cpp:309-345 */
void oracle_hex8_laplacian_apply(ptrdiff_t ne, const int32_t *el, const float *fff, const float *coeff,
                                 const double *x, double *y) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double f[6], K[64], u[8];
        int32_t ev[8];
        const double c = coef(coeff, e);
        for (int k = 0; k < 6; k++) f[k] = c * (double)fff[e * 6 + k];
        oracle_hex8_laplacian_matrix(f, K);
        for (int v = 0; v < 8; v++) {
            ev[v] = el[(size_t)v * ne + e];
            u[v] = x[ev[v]];
        }
        for (int i = 0; i < 8; i++) {
            double s = 0;
            for (int j = 0; j < 8; j++) s += K[i * 8 + j] * u[j];
#pragma omp atomic update
            y[ev[i]] += s;
        }
    }
}

void oracle_hex8_laplacian_diag(ptrdiff_t ne, const int32_t *el, const float *fff, const float *coeff, double *d) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double f[6], K[64];
        const double c = coef(coeff, e);
        for (int k = 0; k < 6; k++) f[k] = c * (double)fff[e * 6 + k];
        oracle_hex8_laplacian_matrix(f, K);
        for (int i = 0; i < 8; i++) {
#pragma omp atomic update
            d[el[(size_t)i * ne + e]] += K[i * 9];
        }
    }
}

/* micro-element corners inside a macro element, hex8 corner order (sshex8.hpp:53-62) */
static inline void micro_corners(int L, int xi, int yi, int zi, int *lev) {
    for (int v = 0; v < 8; v++) lev[v] = lidx(L, xi + hex8_cx[v], yi + hex8_cy[v], zi + hex8_cz[v]);
}

/* operators/sshex8/sshex8_laplacian.cpp:698-769: sub_fff = h * fff (:19-28), every
 * micro element applies the same 8x8 matrix; the reference's stencil+skeleton split
 * is an evaluation order of the same sum. */
void oracle_sshex8_laplacian_apply(int L, ptrdiff_t ne, const int32_t *el, const float *fff, const float *coeff,
                                   const double *x, double *y) {
    const int nxe = oracle_sshex8_nxe(L);
#pragma omp parallel
    {
        double *eu = (double *)malloc(sizeof(double) * nxe);
        double *ev = (double *)malloc(sizeof(double) * nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            const double h = 1.0 / L, c = coef(coeff, e);
            double f[6], K[64];
            for (int k = 0; k < 6; k++) f[k] = (double)fff[e * 6 + k] * h * c;
            oracle_hex8_laplacian_matrix(f, K);
            for (int d = 0; d < nxe; d++) {
                eu[d] = x[el[(size_t)d * ne + e]];
                ev[d] = 0;
            }
            for (int zi = 0; zi < L; zi++)
                for (int yi = 0; yi < L; yi++)
                    for (int xi = 0; xi < L; xi++) {
                        int lev[8];
                        micro_corners(L, xi, yi, zi, lev);
                        for (int i = 0; i < 8; i++) {
                            double s = 0;
                            for (int j = 0; j < 8; j++) s += K[i * 8 + j] * eu[lev[j]];
                            ev[lev[i]] += s;
                        }
                    }
            for (int d = 0; d < nxe; d++) {
#pragma omp atomic update
                y[el[(size_t)d * ne + e]] += ev[d];
            }
        }
        free(eu);
        free(ev);
    }
}

/* Stored sub-element matrices applied to every micro element ("ss:em"):
 * operators/stencil/sshex8_stencil_element_matrix_apply.cpp:10-78 (bs 1, [E][64]) and :80-160
 * (bs 3, [E][576], index component*8 + node; vectors interleaved). */
void oracle_sshex8_em_apply(int L, ptrdiff_t ne, const int32_t *el, int bs, const double *mats, const double *x,
                            double *y) {
    const int nxe = oracle_sshex8_nxe(L), nd = 8 * bs;
#pragma omp parallel
    {
        double *eu = (double *)malloc(sizeof(double) * bs * nxe);
        double *ev = (double *)malloc(sizeof(double) * bs * nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            const double *K = mats + (size_t)e * nd * nd;
            for (int d = 0; d < nxe; d++) {
                const int32_t n = el[(size_t)d * ne + e];
                for (int c = 0; c < bs; c++) {
                    eu[c * nxe + d] = x[(size_t)n * bs + c];
                    ev[c * nxe + d] = 0;
                }
            }
            for (int zi = 0; zi < L; zi++)
                for (int yi = 0; yi < L; yi++)
                    for (int xi = 0; xi < L; xi++) {
                        int lev[8];
                        double u[24];
                        micro_corners(L, xi, yi, zi, lev);
                        for (int c = 0; c < bs; c++)
                            for (int v = 0; v < 8; v++) u[c * 8 + v] = eu[c * nxe + lev[v]];
                        for (int i = 0; i < nd; i++) {
                            double acc = 0;
                            for (int j = 0; j < nd; j++) acc += K[i * nd + j] * u[j];
                            ev[(i / 8) * nxe + lev[i % 8]] += acc;
                        }
                    }
            for (int d = 0; d < nxe; d++) {
                const int32_t n = el[(size_t)d * ne + e];
                for (int c = 0; c < bs; c++) {
#pragma omp atomic update
                    y[(size_t)n * bs + c] += ev[c * nxe + d];
                }
            }
        }
        free(eu);
        free(ev);
    }
}

/* sshex8_laplacian_element_matrix (operators/sshex8/sshex8_laplacian.cpp:595-634) from the macro
 * FFF words: sub-FFF = h * FFF (times the coefficient), then the 8x8 exact-integral matrix. */
void oracle_sshex8_laplacian_element_matrices(int L, ptrdiff_t ne, const float *fff, const float *coeff,
                                              double *out) {
    for (ptrdiff_t e = 0; e < ne; e++) {
        const double h = 1.0 / L, c = coef(coeff, e);
        double f[6];
        for (int k = 0; k < 6; k++) f[k] = (double)fff[e * 6 + k] * h * c;
        oracle_hex8_laplacian_matrix(f, out + (size_t)e * 64);
    }
}

/* operators/sshex8/sshex8_laplacian.cpp:333-435 */
void oracle_sshex8_laplacian_diag(int L, ptrdiff_t ne, const int32_t *el, const float *fff, const float *coeff,
                                  double *d) {
    const int nxe = oracle_sshex8_nxe(L);
#pragma omp parallel
    {
        double *ev = (double *)malloc(sizeof(double) * nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            const double h = 1.0 / L, c = coef(coeff, e);
            double f[6], K[64];
            for (int k = 0; k < 6; k++) f[k] = (double)fff[e * 6 + k] * h * c;
            oracle_hex8_laplacian_matrix(f, K);
            for (int i = 0; i < nxe; i++) ev[i] = 0;
            for (int zi = 0; zi < L; zi++)
                for (int yi = 0; yi < L; yi++)
                    for (int xi = 0; xi < L; xi++) {
                        int lev[8];
                        micro_corners(L, xi, yi, zi, lev);
                        for (int i = 0; i < 8; i++) ev[lev[i]] += K[i * 9];
                    }
            for (int i = 0; i < nxe; i++) {
#pragma omp atomic update
                d[el[(size_t)i * ne + e]] += ev[i];
            }
        }
        free(ev);
    }
}

static void elasticity_matrix_from_words(const float *adj, const float *det, ptrdiff_t e, double mu, double lambda,
                                         double c, double h, double *K) {
    double A[9];
    for (int i = 0; i < 9; i++) A[i] = (double)adj[e * 9 + i] * (h * h);
    const double dt = (double)det[e] * (h * h * h);
    oracle_hex8_elasticity_matrix(mu * c, lambda * c, A, dt, K);
}

/* The sshex8 elasticity sub-element matrices (adjugate * h^2, det * h^3; component*8 + node). */
void oracle_sshex8_elasticity_element_matrices(int L, ptrdiff_t ne, const float *adj, const float *det, double mu,
                                               double lambda, const float *coeff, double *out) {
    for (ptrdiff_t e = 0; e < ne; e++)
        elasticity_matrix_from_words(adj, det, e, mu, lambda, coef(coeff, e), 1.0 / L, out + (size_t)e * 576);
}


/* operators/hex8/hex8_linear_elasticity.cpp:625-760 (affine, adjugate at centre) */
void oracle_hex8_elasticity_apply(ptrdiff_t ne, const int32_t *el, const float *adj, const float *det, double mu,
                                  double lambda, const float *coeff, const double *x, double *y) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double K[576], u[24];
        int32_t ev[8];
        elasticity_matrix_from_words(adj, det, e, mu, lambda, coef(coeff, e), 1.0, K);
        for (int v = 0; v < 8; v++) {
            ev[v] = el[(size_t)v * ne + e];
            for (int c = 0; c < 3; c++) u[c * 8 + v] = x[(size_t)ev[v] * 3 + c];
        }
        for (int i = 0; i < 24; i++) {
            double s = 0;
            for (int j = 0; j < 24; j++) s += K[i * 24 + j] * u[j];
            const int c = i / 8, v = i % 8;
#pragma omp atomic update
            y[(size_t)ev[v] * 3 + c] += s;
        }
    }
}

void oracle_hex8_elasticity_diag(ptrdiff_t ne, const int32_t *el, const float *adj, const float *det, double mu,
                                 double lambda, const float *coeff, double *d) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        double K[576];
        elasticity_matrix_from_words(adj, det, e, mu, lambda, coef(coeff, e), 1.0, K);
        for (int i = 0; i < 24; i++) {
            const int c = i / 8, v = i % 8;
#pragma omp atomic update
            d[(size_t)el[(size_t)v * ne + e] * 3 + c] += K[i * 25];
        }
    }
}

/* operators/sshex8/sshex8_linear_elasticity.cpp:275-425 (sub adjugate * h^2, det * h^3:
 * operators/sshex8/sshex8.hpp:26-42) */
void oracle_sshex8_elasticity_apply(int L, ptrdiff_t ne, const int32_t *el, const float *adj, const float *det,
                                    double mu, double lambda, const float *coeff, const double *x, double *y) {
    const int nxe = oracle_sshex8_nxe(L);
#pragma omp parallel
    {
        double *eu = (double *)malloc(sizeof(double) * 3 * nxe);
        double *ev = (double *)malloc(sizeof(double) * 3 * nxe);
        double *K = (double *)malloc(sizeof(double) * 576);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            elasticity_matrix_from_words(adj, det, e, mu, lambda, coef(coeff, e), 1.0 / L, K);
            for (int d = 0; d < nxe; d++) {
                const int32_t n = el[(size_t)d * ne + e];
                for (int c = 0; c < 3; c++) {
                    eu[c * nxe + d] = x[(size_t)n * 3 + c];
                    ev[c * nxe + d] = 0;
                }
            }
            for (int zi = 0; zi < L; zi++)
                for (int yi = 0; yi < L; yi++)
                    for (int xi = 0; xi < L; xi++) {
                        int lev[8];
                        double u[24];
                        micro_corners(L, xi, yi, zi, lev);
                        for (int c = 0; c < 3; c++)
                            for (int v = 0; v < 8; v++) u[c * 8 + v] = eu[c * nxe + lev[v]];
                        for (int i = 0; i < 24; i++) {
                            double s = 0;
                            for (int j = 0; j < 24; j++) s += K[i * 24 + j] * u[j];
                            ev[(i / 8) * nxe + lev[i % 8]] += s;
                        }
                    }
            for (int d = 0; d < nxe; d++) {
                const int32_t n = el[(size_t)d * ne + e];
                for (int c = 0; c < 3; c++) {
#pragma omp atomic update
                    y[(size_t)n * 3 + c] += ev[c * nxe + d];
                }
            }
        }
        free(eu);
        free(ev);
        free(K);
    }
}

static void sshex8_elasticity_blocks(int L, ptrdiff_t ne, const int32_t *el, const float *adj, const float *det,
                                     double mu, double lambda, const float *coeff, double *d, int full_block) {
    const int nxe = oracle_sshex8_nxe(L);
#pragma omp parallel
    {
        double *acc = (double *)malloc(sizeof(double) * 6 * nxe);
        double *K = (double *)malloc(sizeof(double) * 576);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            elasticity_matrix_from_words(adj, det, e, mu, lambda, coef(coeff, e), 1.0 / L, K);
            for (int i = 0; i < 6 * nxe; i++) acc[i] = 0;
            for (int zi = 0; zi < L; zi++)
                for (int yi = 0; yi < L; yi++)
                    for (int xi = 0; xi < L; xi++) {
                        int lev[8];
                        micro_corners(L, xi, yi, zi, lev);
                        for (int v = 0; v < 8; v++) {
                            /* (xx, xy, xz, yy, yz, zz) */
                            static const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
                            for (int k = 0; k < 6; k++)
                                acc[lev[v] * 6 + k] += K[(ca[k] * 8 + v) * 24 + cb[k] * 8 + v];
                        }
                    }
            for (int i = 0; i < nxe; i++) {
                const int32_t n = el[(size_t)i * ne + e];
                if (full_block) {
                    for (int k = 0; k < 6; k++) {
#pragma omp atomic update
                        d[(size_t)n * 6 + k] += acc[i * 6 + k];
                    }
                } else {
                    static const int dk[3] = {0, 3, 5};
                    for (int c = 0; c < 3; c++) {
#pragma omp atomic update
                        d[(size_t)n * 3 + c] += acc[i * 6 + dk[c]];
                    }
                }
            }
        }
        free(acc);
        free(K);
    }
}

void oracle_sshex8_elasticity_diag(int L, ptrdiff_t ne, const int32_t *el, const float *adj, const float *det,
                                   double mu, double lambda, const float *coeff, double *d) {
    sshex8_elasticity_blocks(L, ne, el, adj, det, mu, lambda, coeff, d, 0);
}

void oracle_sshex8_elasticity_block_diag(int L, ptrdiff_t ne, const int32_t *el, const float *adj, const float *det,
                                         double mu, double lambda, const float *coeff, double *d) {
    sshex8_elasticity_blocks(L, ne, el, adj, det, mu, lambda, coeff, d, 1);
}

/* =====================================================================
 * Transfers
 * ===================================================================== */

void oracle_sshex8_prolongate(ptrdiff_t ne, int from_level, const int32_t *from_el, int to_level,
                              const int32_t *to_el, int vs, const double *from, double *to) {
    const int s = to_level / from_level;
    const int to_nxe = oracle_sshex8_nxe(to_level);
#pragma omp parallel
    {
        double *c = (double *)malloc(sizeof(double) * to_nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            for (int d = 0; d < vs; d++) {
                /* coarse nodes copied, then every fine node is the trilinear
                 * interpolant of its coarse cell's corners -- the x/y/z edge,
                 * face and volume passes of :313-606 compose to exactly this. */
                for (int z = 0; z <= to_level; z++)
                    for (int y = 0; y <= to_level; y++)
                        for (int x = 0; x <= to_level; x++) {
                            int cx = x / s, cy = y / s, cz = z / s;
                            if (cx == from_level) cx--;
                            if (cy == from_level) cy--;
                            if (cz == from_level) cz--;
                            const double tx = (double)(x - cx * s) / s, ty = (double)(y - cy * s) / s,
                                         tz = (double)(z - cz * s) / s;
                            double v = 0;
                            for (int k = 0; k < 8; k++) {
                                const double w = (hex8_cx[k] ? tx : 1 - tx) * (hex8_cy[k] ? ty : 1 - ty) *
                                                 (hex8_cz[k] ? tz : 1 - tz);
                                if (w == 0) continue;
                                const int32_t n = from_el[(size_t)lidx(from_level, cx + hex8_cx[k], cy + hex8_cy[k],
                                                                      cz + hex8_cz[k]) * ne + e];
                                v += w * from[(size_t)n * vs + d];
                            }
                            c[lidx(to_level, x, y, z)] = v;
                        }
                for (int i = 0; i < to_nxe; i++) to[(size_t)to_el[(size_t)i * ne + e] * vs + d] = c[i];
            }
        }
        free(c);
    }
}

void oracle_sshex8_restrict(ptrdiff_t ne, int from_level, const int32_t *from_el, const uint16_t *count, int to_level,
                            const int32_t *to_el, int vs, const double *from, double *to) {
    const int s = from_level / to_level;
    const int to_nxe = oracle_sshex8_nxe(to_level);
#pragma omp parallel
    {
        double *out = (double *)malloc(sizeof(double) * to_nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            for (int d = 0; d < vs; d++) {
                for (int i = 0; i < to_nxe; i++) out[i] = 0;
                for (int z = 0; z <= from_level; z++)
                    for (int y = 0; y <= from_level; y++)
                        for (int x = 0; x <= from_level; x++) {
                            const int32_t gid = from_el[(size_t)lidx(from_level, x, y, z) * ne + e];
                            const double val = from[(size_t)gid * vs + d] / count[gid];
                            const int cx = x / s, cy = y / s, cz = z / s;
                            const int xinc = (cx + 1 < to_level ? cx + 1 : to_level) - cx;
                            const int yinc = (cy + 1 < to_level ? cy + 1 : to_level) - cy;
                            const int zinc = (cz + 1 < to_level ? cz + 1 : to_level) - cz;
                            const double lx = (double)(x - cx * s) / s, ly = (double)(y - cy * s) / s,
                                         lz = (double)(z - cz * s) / s;
                            const double px[2] = {1 - lx, lx}, py[2] = {1 - ly, ly}, pz[2] = {1 - lz, lz};
                            for (int kk = 0; kk <= zinc; kk++)
                                for (int jj = 0; jj <= yinc; jj++)
                                    for (int ii = 0; ii <= xinc; ii++)
                                        out[lidx(to_level, cx + ii, cy + jj, cz + kk)] += px[ii] * py[jj] * pz[kk] * val;
                        }
                for (int i = 0; i < to_nxe; i++) {
#pragma omp atomic update
                    to[(size_t)to_el[(size_t)i * ne + e] * vs + d] += out[i];
                }
            }
        }
        free(out);
    }
}

void oracle_sshex8_hierarchical_prolongation(int L, ptrdiff_t ne, const int32_t *el, int vs, const double *from,
                                             double *to) {
    const double h = 1.0 / L;
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        int32_t cn[8];
        for (int v = 0; v < 8; v++)
            cn[v] = el[(size_t)lidx(L, hex8_cx[v] * L, hex8_cy[v] * L, hex8_cz[v] * L) * ne + e];
        for (int z = 0; z <= L; z++)
            for (int y = 0; y <= L; y++)
                for (int x = 0; x <= L; x++) {
                    const double qx = x * h, qy = y * h, qz = z * h;
                    const int32_t n = el[(size_t)lidx(L, x, y, z) * ne + e];
                    for (int d = 0; d < vs; d++) {
                        double v = 0;
                        for (int k = 0; k < 8; k++)
                            v += (hex8_cx[k] ? qx : 1 - qx) * (hex8_cy[k] ? qy : 1 - qy) * (hex8_cz[k] ? qz : 1 - qz) *
                                 from[(size_t)cn[k] * vs + d];
                        to[(size_t)n * vs + d] = v;
                    }
                }
    }
}

void oracle_sshex8_hierarchical_restriction(int L, ptrdiff_t ne, const int32_t *el, const uint16_t *count, int vs,
                                            const double *from, double *to) {
    const double h = 1.0 / L;
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        int32_t cn[8];
        for (int v = 0; v < 8; v++)
            cn[v] = el[(size_t)lidx(L, hex8_cx[v] * L, hex8_cy[v] * L, hex8_cz[v] * L) * ne + e];
        for (int d = 0; d < vs; d++) {
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int z = 0; z <= L; z++)
                for (int y = 0; y <= L; y++)
                    for (int x = 0; x <= L; x++) {
                        const double qx = x * h, qy = y * h, qz = z * h;
                        const int32_t n = el[(size_t)lidx(L, x, y, z) * ne + e];
                        const double val = from[(size_t)n * vs + d] / count[n];
                        for (int k = 0; k < 8; k++)
                            acc[k] += (hex8_cx[k] ? qx : 1 - qx) * (hex8_cy[k] ? qy : 1 - qy) *
                                      (hex8_cz[k] ? qz : 1 - qz) * val;
                    }
            for (int k = 0; k < 8; k++) {
#pragma omp atomic update
                to[(size_t)cn[k] * vs + d] += acc[k];
            }
        }
    }
}

/* =====================================================================
 * Dirichlet
 * ===================================================================== */

void oracle_constraint_nodes_to_value_vec(ptrdiff_t n, const int32_t *nodes, int bs, int comp, double value,
                                          double *x) {
    for (ptrdiff_t i = 0; i < n; i++) x[(size_t)nodes[i] * bs + comp] = value;
}

void oracle_constraint_nodes_copy_vec(ptrdiff_t n, const int32_t *nodes, int bs, int comp, const double *src,
                                      double *dst) {
    for (ptrdiff_t i = 0; i < n; i++) dst[(size_t)nodes[i] * bs + comp] = src[(size_t)nodes[i] * bs + comp];
}
