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

/* Standard hex8 corner -> structured coordinate: the corner order of box_mesh.py:520-534 and
 * operators/sshex8/sshex8.hpp:53-62 (lagr_to_proteus_corners, mesh/sfem_hex8_mesh_graph.cpp:451-459). */
static const int hex8_cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int hex8_cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int hex8_cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

/* hex8_edge_connectivity, LAGRANGE order (mesh/sfem_hex8_mesh_graph.cpp:37-47 and the local
 * copy lagr_connectivity at :534-543): the three corners joined to each corner by an edge. */
static const int hex8_edge_conn[8][3] = {{1, 3, 4}, {0, 2, 5}, {1, 3, 6}, {0, 2, 7},
                                         {0, 5, 7}, {1, 4, 6}, {2, 5, 7}, {3, 4, 6}};

/* HEX8 local side table (4 corners per side, cyclic).  The reference fills it with smesh's
 * fill_local_side_table (mesh/sfem_hex8_mesh_graph.cpp:652), which is absent; this is the
 * in-tree HEX8 side list of python/codegen/hex8.py:44-52 (exodus side order). */
static const int hex8_sides[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                     {3, 0, 4, 7}, {3, 2, 1, 0}, {4, 5, 6, 7}};

static int cmp_i32(const void *a, const void *b) {
    const int32_t p = *(const int32_t *)a, q = *(const int32_t *)b;
    return (p > q) - (p < q);
}

/* sortreduce: sort ascending and drop duplicates; returns the new length. */
static int sortreduce_i32(int32_t *v, int n) {
    if (n == 0) return 0;
    qsort(v, (size_t)n, sizeof(int32_t), cmp_i32);
    int m = 1;
    for (int i = 1; i < n; i++)
        if (v[i] != v[m - 1]) v[m++] = v[i];
    return m;
}

/* node -> elements (CSR), elements in ascending order */
static void build_n2e(ptrdiff_t ne, ptrdiff_t nn, const int32_t *me, ptrdiff_t **ptr_out, ptrdiff_t **idx_out) {
    ptrdiff_t *ptr = (ptrdiff_t *)calloc((size_t)nn + 1, sizeof(ptrdiff_t));
    for (int v = 0; v < 8; v++)
        for (ptrdiff_t e = 0; e < ne; e++) ptr[me[v * ne + e] + 1]++;
    for (ptrdiff_t i = 0; i < nn; i++) ptr[i + 1] += ptr[i];
    ptrdiff_t *fill = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * ((size_t)nn + 1));
    memcpy(fill, ptr, sizeof(ptrdiff_t) * ((size_t)nn + 1));
    ptrdiff_t *idx = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * (size_t)(ptr[nn] + 1));
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int v = 0; v < 8; v++) {
            const int32_t n = me[v * ne + e];
            /* an element listing a node twice is degenerate; keep one entry */
            int dup = 0;
            for (int w = 0; w < v; w++) dup |= me[w * ne + e] == n;
            if (!dup) idx[fill[n]++] = e;
        }
    /* entries were appended in element order: already ascending; fix the pointers for dups */
    ptrdiff_t w = 0;
    for (ptrdiff_t i = 0; i < nn; i++) {
        const ptrdiff_t b = ptr[i], end = fill[i];
        ptr[i] = w;
        for (ptrdiff_t k = b; k < end; k++) idx[w++] = idx[k];
    }
    ptr[nn] = w;
    free(fill);
    *ptr_out = ptr;
    *idx_out = idx;
}

/* hex8_build_edge_graph_from_n2e (mesh/sfem_hex8_mesh_graph.cpp:52-135): row i holds the
 * sorted unique corners joined to node i by an element edge. */
static void hex8_edge_graph(ptrdiff_t ne, ptrdiff_t nn, const int32_t *me, const ptrdiff_t *n2eptr,
                            const ptrdiff_t *n2e, ptrdiff_t **rowptr_out, int32_t **colidx_out) {
    ptrdiff_t *rowptr = (ptrdiff_t *)calloc((size_t)nn + 1, sizeof(ptrdiff_t));
    ptrdiff_t maxdeg = 0;
    for (ptrdiff_t i = 0; i < nn; i++)
        if (n2eptr[i + 1] - n2eptr[i] > maxdeg) maxdeg = n2eptr[i + 1] - n2eptr[i];
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(3 * maxdeg + 1));
    for (int pass = 0; pass < 2; pass++) {
        int32_t *colidx = pass ? *colidx_out : NULL;
        for (ptrdiff_t node = 0; node < nn; node++) {
            int nb = 0;
            for (ptrdiff_t k = n2eptr[node]; k < n2eptr[node + 1]; k++) {
                const ptrdiff_t e = n2e[k];
                int l = 0;
                for (int v = 0; v < 8; v++)
                    if (me[v * ne + e] == node) {
                        l = v;
                        break;
                    }
                for (int d = 0; d < 3; d++) buf[nb++] = me[hex8_edge_conn[l][d] * ne + e];
            }
            nb = sortreduce_i32(buf, nb);
            if (!pass)
                rowptr[node + 1] = nb;
            else
                memcpy(&colidx[rowptr[node]], buf, sizeof(int32_t) * (size_t)nb);
        }
        if (!pass) {
            for (ptrdiff_t i = 0; i < nn; i++) rowptr[i + 1] += rowptr[i];
            *colidx_out = (int32_t *)malloc(sizeof(int32_t) * (size_t)(rowptr[nn] + 1));
        }
    }
    free(buf);
    *rowptr_out = rowptr;
}

/* element adjacency through sides (the role of smesh's create_element_adj_table,
 * mesh/sfem_hex8_mesh_graph.cpp:655): adj[e*6+f] = the other element holding the four corners of
 * side f of e, or -1.  On a conforming hex8 mesh a side has at most two elements. */
static void hex8_adj_table(ptrdiff_t ne, const int32_t *me, const ptrdiff_t *n2eptr, const ptrdiff_t *n2e,
                           ptrdiff_t *adj) {
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int f = 0; f < 6; f++) {
            int32_t s[4];
            for (int i = 0; i < 4; i++) s[i] = me[hex8_sides[f][i] * ne + e];
            adj[e * 6 + f] = -1;
            for (ptrdiff_t k = n2eptr[s[0]]; k < n2eptr[s[0] + 1]; k++) {
                const ptrdiff_t o = n2e[k];
                if (o == e) continue;
                int hits = 0;
                for (int i = 1; i < 4; i++)
                    for (int v = 0; v < 8; v++)
                        if (me[v * ne + o] == s[i]) {
                            hits++;
                            break;
                        }
                if (hits == 3) {
                    adj[e * 6 + f] = o;
                    break;
                }
            }
        }
}

/* index_face (mesh/sfem_hex8_mesh_graph.cpp:346-446): numbers the interior nodes of side f of
 * element e.  o = the side corner with the smallest global id, u and v its two neighbours on the
 * side ordered by global id; nodes run u fastest, then v, starting next to o. */
static void index_face(int L, ptrdiff_t ne, const int32_t *me, ptrdiff_t face_offset, ptrdiff_t e, int f,
                       int32_t *elements) {
    int argmin = 0;
    int32_t valmin = me[hex8_sides[f][0] * ne + e];
    for (int i = 0; i < 4; i++) {
        const int32_t t = me[hex8_sides[f][i] * ne + e];
        if (t < valmin) {
            argmin = i;
            valmin = t;
        }
    }
    const int lst_o = argmin;
    int lst_u = (lst_o + 1) % 4, lst_v = (lst_o + 3) % 4;
    if (me[hex8_sides[f][lst_u] * ne + e] > me[hex8_sides[f][lst_v] * ne + e]) {
        const int t = lst_v;
        lst_v = lst_u;
        lst_u = t;
    }
    const int co = hex8_sides[f][lst_o], cu = hex8_sides[f][lst_u], cv = hex8_sides[f][lst_v];
    const int oc[3] = {hex8_cx[co] * L, hex8_cy[co] * L, hex8_cz[co] * L};
    const int uc[3] = {hex8_cx[cu] * L, hex8_cy[cu] * L, hex8_cz[cu] * L};
    const int vc[3] = {hex8_cx[cv] * L, hex8_cy[cv] * L, hex8_cz[cv] * L};
    int o_start[3], u_len[3], u_dir[3], v_len[3], v_dir[3];
    for (int d = 0; d < 3; d++) o_start[d] = oc[d];
    for (int d = 0; d < 3; d++) {
        int x = uc[d] - oc[d];
        u_dir[d] = 1;
        u_len[d] = 1;
        if (x > 0) {
            u_len[d] = x - 1;
            o_start[d] = 1;
        } else if (x < 0) {
            u_len[d] = x + 1;
            u_dir[d] = -1;
            o_start[d] = L - 1;
        }
    }
    for (int d = 0; d < 3; d++) {
        int x = vc[d] - oc[d];
        v_dir[d] = 1;
        v_len[d] = 1;
        if (x > 0) {
            v_len[d] = x - 1;
            o_start[d] = 1;
        } else if (x < 0) {
            v_len[d] = x + 1;
            v_dir[d] = -1;
            o_start[d] = L - 1;
        }
    }
    ptrdiff_t local = 0;
    for (int vz = 0; vz != v_len[2]; vz += v_dir[2])
        for (int vy = 0; vy != v_len[1]; vy += v_dir[1])
            for (int vx = 0; vx != v_len[0]; vx += v_dir[0])
                for (int uz = 0; uz != u_len[2]; uz += u_dir[2])
                    for (int uy = 0; uy != u_len[1]; uy += u_dir[1])
                        for (int ux = 0; ux != u_len[0]; ux += u_dir[0]) {
                            const int p = lidx(L, ux + vx + o_start[0], uy + vy + o_start[1], uz + vz + o_start[2]);
                            elements[(size_t)p * ne + e] = (int32_t)(face_offset + local++);
                        }
}

/*
 * sshex8_generate_elements (mesh/sfem_hex8_mesh_graph.cpp:448-735, commented out in the
 * checkout; the live one is smesh::to_semistructured, absent).  Raw numbering:
 *   1) corners keep their macro ids;
 *   2) edge-interior nodes: index_base + edge_id*(L-1) + en, where edge ids enumerate the pairs
 *      (i, j), i < j, of the hex8 edge graph in CRS order (rows ascending, sorted columns), and
 *      en runs from the lower-id corner towards the higher-id one (:510-633);
 *   3) face-interior nodes: sides visited element by element, side by side; a side already
 *      reached from a lower-numbered neighbour is skipped; both elements of a side are filled
 *      with index_face (:636-697);
 *   4) volume-interior nodes: base + e*(L-1)^3 + local, x fastest (:702-722).
 * elements: SoA [(L+1)^3][ne].  Returns the node count (interior_start + ne*(L-1)^3).
 */
static ptrdiff_t sshex8_generate_elements(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me,
                                          int32_t *elements) {
    ptrdiff_t *n2eptr, *n2e;
    build_n2e(ne, nnodes, me, &n2eptr, &n2e);

    for (int d = 0; d < 8; d++) {
        const int p = lidx(L, hex8_cx[d] * L, hex8_cy[d] * L, hex8_cz[d] * L);
        for (ptrdiff_t e = 0; e < ne; e++) elements[(size_t)p * ne + e] = me[d * ne + e];
    }
    ptrdiff_t index_base = nnodes;

    const int nxedge = L - 1;
    if (nxedge > 0) {
        ptrdiff_t *rowptr;
        int32_t *colidx = NULL;
        hex8_edge_graph(ne, nnodes, me, n2eptr, n2e, &rowptr, &colidx);
        const ptrdiff_t nnz = rowptr[nnodes];
        int32_t *edge_idx = (int32_t *)calloc((size_t)nnz + 1, sizeof(int32_t));
        int32_t next_id = 0;
        for (ptrdiff_t i = 0; i < nnodes; i++)
            for (ptrdiff_t k = rowptr[i]; k < rowptr[i + 1]; k++)
                if (i < colidx[k]) edge_idx[k] = next_id++;
        const ptrdiff_t nedges = next_id;

        for (ptrdiff_t e = 0; e < ne; e++) {
            int32_t nodes[8];
            for (int d = 0; d < 8; d++) nodes[d] = me[d * ne + e];
            for (int d1 = 0; d1 < 8; d1++) {
                const int32_t node1 = nodes[d1];
                for (int d2 = 0; d2 < 3; d2++) {
                    const int c2 = hex8_edge_conn[d1][d2];
                    const int32_t node2 = nodes[c2];
                    if (node1 > node2) continue; /* edges run from the smaller id */
                    ptrdiff_t k = rowptr[node1];
                    while (colidx[k] != node2) k++;
                    const ptrdiff_t g_edge = edge_idx[k];
                    const int c1[3] = {hex8_cx[d1] * L, hex8_cy[d1] * L, hex8_cz[d1] * L};
                    const int cc[3] = {hex8_cx[c2] * L, hex8_cy[c2] * L, hex8_cz[c2] * L};
                    int start[3], len[3], dir[3];
                    for (int d = 0; d < 3; d++) {
                        start[d] = c1[d];
                        int x = cc[d] - c1[d];
                        dir[d] = 1;
                        len[d] = 1;
                        if (x > 0) {
                            len[d] = x - 1;
                            start[d] = 1;
                        } else if (x < 0) {
                            len[d] = x + 1;
                            dir[d] = -1;
                            start[d] = L - 1;
                        }
                    }
                    const ptrdiff_t edge_start = index_base + g_edge * nxedge;
                    int en = 0;
                    for (int zi = 0; zi != len[2]; zi += dir[2])
                        for (int yi = 0; yi != len[1]; yi += dir[1])
                            for (int xi = 0; xi != len[0]; xi += dir[0]) {
                                const int p = lidx(L, start[0] + xi, start[1] + yi, start[2] + zi);
                                elements[(size_t)p * ne + e] = (int32_t)(edge_start + en++);
                            }
                }
            }
        }
        free(edge_idx);
        free(rowptr);
        free(colidx);
        index_base += nedges * nxedge;
    }

    const ptrdiff_t nxf = (ptrdiff_t)(L - 1) * (L - 1);
    if (nxf > 0) {
        ptrdiff_t *adj = (ptrdiff_t *)malloc(sizeof(ptrdiff_t) * 6 * (size_t)ne + 1);
        hex8_adj_table(ne, me, n2eptr, n2e, adj);
        ptrdiff_t n_unique_faces = 0;
        for (ptrdiff_t e = 0; e < ne; e++)
            for (int f = 0; f < 6; f++) {
                const ptrdiff_t nb = adj[e * 6 + f];
                if (nb >= 0 && nb < e) continue;
                const ptrdiff_t offset = index_base + n_unique_faces * nxf;
                index_face(L, ne, me, offset, e, f, elements);
                if (nb >= 0) {
                    int nf = 0;
                    for (; nf < 6; nf++)
                        if (adj[nb * 6 + nf] == e) break;
                    index_face(L, ne, me, offset, nb, nf, elements);
                }
                n_unique_faces++;
            }
        free(adj);
        index_base += n_unique_faces * nxf;
    }

    const ptrdiff_t nxelement = (ptrdiff_t)(L - 1) * (L - 1) * (L - 1);
    if (nxelement > 0) {
        const int Lm1 = L - 1;
        for (int zi = 1; zi < L; zi++)
            for (int yi = 1; yi < L; yi++)
                for (int xi = 1; xi < L; xi++) {
                    const int p = lidx(L, xi, yi, zi);
                    const ptrdiff_t en = ((ptrdiff_t)(zi - 1) * Lm1 + (yi - 1)) * Lm1 + xi - 1;
                    for (ptrdiff_t e = 0; e < ne; e++)
                        elements[(size_t)p * ne + e] = (int32_t)(index_base + e * nxelement + en);
                }
    }
    free(n2eptr);
    free(n2e);
    return index_base + ne * nxelement;
}

/* sshex8_hierarchical_n_levels / sshex8_hierarchical_mesh_levels
 * (mesh/sfem_hex8_mesh_graph.cpp:964-992): 1, then the even chain ... L/4, L/2, L that ends at L
 * (8 -> [1, 2, 4, 8]; 12 -> [1, 6, 12]; 6 -> [1, 6]).  For odd L > 1 the reference list is [1]
 * and its renumbering stops with "Uninitialized node mapping"; this restatement appends L
 * (extension: [1, L]), which changes nothing where the reference succeeds. */
int oracle_sshex8_hierarchical_levels(int L, int *levels) {
    int count = 0;
    int l = L;
    for (; l > 1 && l % 2 == 0; l /= 2) count++;
    if (l >= 1) count++;
    int n = count;
    int c = 0;
    l = L;
    for (; l > 1 && l % 2 == 0; l /= 2) {
        if (levels) levels[n - 1 - c] = l;
        c++;
    }
    if (levels) levels[0] = 1;
    if (n == 1 && L > 1) {
        if (levels) levels[1] = L;
        n = 2;
    }
    return n;
}

/* sshex8_hierarchical_renumbering (mesh/sfem_hex8_mesh_graph.cpp:994-1060): macro corners keep
 * their ids; then, level by level from the coarsest (levels[1]) to L, elements in order and the
 * level's nodes of each element in (z, y, x) order, every node not yet numbered takes the next id.
 * Every derefined level is therefore a prefix of the numbering.  In place on elements [nxe][ne];
 * returns the node count. */
static ptrdiff_t sshex8_hierarchical_renumbering(int L, ptrdiff_t ne, ptrdiff_t nnodes, int32_t *elements) {
    int levels[32];
    const int nlevels = oracle_sshex8_hierarchical_levels(L, levels);
    int32_t *map = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nnodes + 1));
    for (ptrdiff_t i = 0; i < nnodes; i++) map[i] = -1;
    int32_t next_id = 0;
    for (int zi = 0; zi <= 1; zi++)
        for (int yi = 0; yi <= 1; yi++)
            for (int xi = 0; xi <= 1; xi++)
                for (ptrdiff_t e = 0; e < ne; e++) {
                    const int32_t idx = elements[(size_t)lidx(L, xi * L, yi * L, zi * L) * ne + e];
                    map[idx] = idx;
                    if (idx > next_id) next_id = idx;
                }
    next_id++;
    for (int k = 1; k < nlevels; k++) {
        const int l = levels[k], sf = L / l;
        for (ptrdiff_t e = 0; e < ne; e++)
            for (int zi = 0; zi <= l; zi++)
                for (int yi = 0; yi <= l; yi++)
                    for (int xi = 0; xi <= l; xi++) {
                        const int32_t idx = elements[(size_t)lidx(L, xi * sf, yi * sf, zi * sf) * ne + e];
                        if (map[idx] == -1) map[idx] = next_id++;
                    }
    }
    const ptrdiff_t nxe = oracle_sshex8_nxe(L);
    for (ptrdiff_t i = 0; i < nxe * ne; i++) elements[i] = map[elements[i]];
    free(map);
    return next_id;
}

ptrdiff_t oracle_sshex8_refine_ex(int L, int hierarchical, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me,
                                  const float *mp, int32_t *elements, float *points) {
    const int nxe = oracle_sshex8_nxe(L);
    int32_t *el = elements ? elements : (int32_t *)malloc(sizeof(int32_t) * (size_t)nxe * ne + 4);
    ptrdiff_t ntotal = sshex8_generate_elements(L, ne, nnodes, me, el);
    if (hierarchical) ntotal = sshex8_hierarchical_renumbering(L, ne, ntotal, el);
    if (!elements) {
        free(el);
        return ntotal;
    }
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
    return ntotal;
}

ptrdiff_t oracle_sshex8_refine(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me, const float *mp,
                               int32_t *elements, float *points) {
    return oracle_sshex8_refine_ex(L, 1, ne, nnodes, me, mp, elements, points);
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

/* operators/tet4/tet4_fff.cpp:5-28 + tet4_fff (operators/tet4/tet4_inline_cpu.hpp:224-278).
 * The reference keeps every temporary in float32 (geom_t) with this evaluation order; two steps
 * are evaluated in double and rounded to float: the volume det/6 (its 1/6 constants are double)
 * and 1/det^2 (the numerator 1. is double).  Reproducing the reference's float32 words needs the
 * same order and no FMA contraction (the oracle is built with -ffp-contract=off).
 * e1, e2, e3: the tet's edge vectors from vertex 0 (the columns of J). */
static void tet4_fff_f32(const float *px, const float *py, const float *pz, float *fff) {
    const float e1x = px[1] - px[0], e2x = px[2] - px[0], e3x = px[3] - px[0];
    const float e1y = py[1] - py[0], e2y = py[2] - py[0], e3y = py[3] - py[0];
    const float e1z = pz[1] - pz[0], e2z = pz[2] - pz[0], e3z = pz[3] - pz[0];
    /* the six products of det(J) */
    const float t_yz = e2y * e3z, p0 = e1x * t_yz;
    const float t_zy = e3y * e2z, p1 = e1x * t_zy;
    const float t_xz = e2x * e3z, p2 = t_xz * e1y;
    const float t_xy = e2x * e3y, p3 = e1z * t_xy;
    const float p4 = e3x * e2z * e1y;
    const float t_yx = e2y * e3x, p5 = e1z * t_yx;
    const double sixth = 1.0 / 6.0;
    const float vol = (float)(-sixth * p2 + sixth * p3 + sixth * p4 - sixth * p5 + sixth * p0 - sixth * p1);
    const float det = -p2 + p3 + p4 - p5 + p0 - p1;
    const float inv_det2 = (float)(1. / (det * det));
    /* cofactors (adjugate entries) of J, row by row */
    const float a00 = t_xy - t_yx, a01 = e3x * e2z - t_xz, a02 = t_yz - t_zy;
    const float a10 = e3x * e1y - e1x * e3y, a11 = e1x * e3z - e1z * e3x, a12 = e1z * e3y - e3z * e1y;
    const float a20 = e1x * e2y - e2x * e1y, a21 = e2x * e1z - e1x * e2z, a22 = e2z * e1y - e2y * e1z;
    const float a00s = a00 * inv_det2, a01s = inv_det2 * a01, a02s = inv_det2 * a02;
    fff[0] = vol * (a00 * a00 * inv_det2 + inv_det2 * (a01 * a01) + inv_det2 * (a02 * a02));
    fff[1] = vol * (a10 * a00s + a11 * a01s + a12 * a02s);
    fff[2] = vol * (a00s * a20 + a01s * a21 + a02s * a22);
    fff[3] = vol * (inv_det2 * (a10 * a10) + inv_det2 * (a11 * a11) + inv_det2 * (a12 * a12));
    fff[4] = vol * (inv_det2 * a10 * a20 + inv_det2 * a11 * a21 + inv_det2 * a12 * a22);
    fff[5] = vol * (inv_det2 * (a20 * a20) + inv_det2 * (a21 * a21) + inv_det2 * (a22 * a22));
}

void oracle_tet4_fff(ptrdiff_t ne, const int32_t *el, ptrdiff_t nn, const float *p, float *fff) {
#pragma omp parallel for
    for (ptrdiff_t e = 0; e < ne; e++) {
        float px[4], py[4], pz[4];
        for (int i = 0; i < 4; i++) {
            const int32_t v = el[(size_t)i * ne + e];
            px[i] = p[0 * nn + v];
            py[i] = p[1 * nn + v];
            pz[i] = p[2 * nn + v];
        }
        tet4_fff_f32(px, py, pz, &fff[e * 6]);
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

/* The reference's timed sshex8 Laplacian path, affine_sshex8_laplacian_stencil_apply_fff
 * (operators/sshex8/sshex8_laplacian.cpp:698-769): per macro element gather its (L+1)^3 values, build
 * the 8x8 sub-element matrix (hex8_sub_fff_0 + hex8_laplacian_matrix_fff_integral), turn it into the
 * 27-point stencil (hex8_matrix_to_stencil, operators/stencil/stencil3.hpp:205-233: the stencil
 * entry of offset d sums A[c][c + d] over the corners c), apply that stencil to the (L-1)^3 interior
 * nodes (sshex8_stencil, :167-184, slice by slice), and give the nodes on the element's faces, edges
 * and corners the contributions of the micro elements present (sshex8_surface_stencil,
 * operators/stencil/sshex8_skeleton_stencil.hpp:1683-: one generated routine per boundary region,
 * restated here as the octant sum of A's row for the node's corner), then scatter with atomics. */
void oracle_sshex8_laplacian_stencil_apply(int L, ptrdiff_t ne, const int32_t *el, const float *fff, const float *coeff,
                                           const double *x, double *y) {
    const int N1 = L + 1, nxe = oracle_sshex8_nxe(L);
#pragma omp parallel
    {
        double *eu = (double *)malloc(sizeof(double) * nxe);
        double *ev = (double *)malloc(sizeof(double) * nxe);
#pragma omp for
        for (ptrdiff_t e = 0; e < ne; e++) {
            const double h = 1.0 / L, c = coef(coeff, e);
            double f[6], K[64], st[27];
            for (int k = 0; k < 6; k++) f[k] = (double)fff[e * 6 + k] * h * c;
            oracle_hex8_laplacian_matrix(f, K);
            for (int k = 0; k < 27; k++) st[k] = 0;
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) {
                    const int dx = hex8_cx[b] - hex8_cx[a], dy = hex8_cy[b] - hex8_cy[a],
                              dz = hex8_cz[b] - hex8_cz[a];
                    st[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] += K[a * 8 + b];
                }
            for (int d = 0; d < nxe; d++) {
                eu[d] = x[el[(size_t)d * ne + e]];
                ev[d] = 0;
            }
            /* interior: the 27-point stencil, slice by slice */
            for (int z = 1; z < L; z++)
                for (int yy = 1; yy < L; yy++)
                    for (int xx = 1; xx < L; xx++) {
                        double acc = 0;
                        for (int dz = -1; dz <= 1; dz++)
                            for (int dy = -1; dy <= 1; dy++)
                                for (int dx = -1; dx <= 1; dx++)
                                    acc += st[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] *
                                           eu[((z + dz) * N1 + yy + dy) * N1 + xx + dx];
                        ev[(z * N1 + yy) * N1 + xx] += acc;
                    }
            /* faces, edges, corners: the micro elements present around the node */
            for (int z = 0; z <= L; z++)
                for (int yy = 0; yy <= L; yy++)
                    for (int xx = 0; xx <= L; xx++) {
                        if (xx > 0 && xx < L && yy > 0 && yy < L && z > 0 && z < L) continue;
                        double acc = 0;
                        for (int a = 0; a < 8; a++) {
                            /* micro element whose corner a is this node */
                            const int x0 = xx - hex8_cx[a], y0 = yy - hex8_cy[a], z0 = z - hex8_cz[a];
                            if (x0 < 0 || x0 >= L || y0 < 0 || y0 >= L || z0 < 0 || z0 >= L) continue;
                            for (int b = 0; b < 8; b++)
                                acc += K[a * 8 + b] *
                                       eu[((z0 + hex8_cz[b]) * N1 + y0 + hex8_cy[b]) * N1 + x0 + hex8_cx[b]];
                        }
                        ev[(z * N1 + yy) * N1 + xx] += acc;
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

/* STREAM triad a = b + s c over n doubles (host bandwidth of the CPU baseline's machine). */
void oracle_stream_triad(ptrdiff_t n, double *a, const double *b, const double *c, double s) {
#pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 0; i < n; i++) a[i] = b[i] + s * c[i];
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
