// Host-side mesh and geometry setup: box meshes, sshex8 refinement with hierarchical
// numbering, FFF and adjugate caches.  (Setup only; everything per-apply runs on the GPU.)
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.hpp"

using namespace sfem_amd;

namespace {

// np.linspace(0, w, num=n, dtype=float32) as box_mesh.py computes it
float linspace_f32(int i, int n, double w) {
    if (n == 1) return 0.f;
    if (i == n - 1) return (float)w;
    return (float)((double)i * (w / (double)(n - 1)));
}

// the 12 hex8 edges as corner pairs
constexpr int kEdges[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                               {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// occurrence -> rank of its key among the unique keys, ordered by key (by_key) or by the first
// occurrence of each key (occurrence = element * slots + slot)
template <typename Key>
std::vector<int64_t> rank_keys(const std::vector<Key> &keys, bool by_key, int64_t &n_unique) {
    const int64_t n = (int64_t)keys.size();
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (keys[a] != keys[b]) return keys[a] < keys[b];
        return a < b;
    });
    std::vector<int64_t> slot(n), first;
    n_unique = 0;
    for (int64_t i = 0; i < n; i++) {
        if (i == 0 || keys[order[i]] != keys[order[i - 1]]) {
            first.push_back(order[i]);
            n_unique++;
        }
        slot[order[i]] = n_unique - 1;
    }
    if (by_key) return slot;
    std::vector<int64_t> by_first(n_unique);
    std::iota(by_first.begin(), by_first.end(), 0);
    std::sort(by_first.begin(), by_first.end(), [&](int64_t a, int64_t b) { return first[a] < first[b]; });
    std::vector<int64_t> rank(n_unique);
    for (int64_t r = 0; r < n_unique; r++) rank[by_first[r]] = r;
    for (int64_t i = 0; i < n; i++) slot[i] = rank[slot[i]];
    return slot;
}

void center_jacobian(const double *x, const double *y, const double *z, double *J) {
#pragma clang fp contract(off)
    const double *c[3] = {x, y, z};
    for (int i = 0; i < 9; i++) J[i] = 0;
    for (int v = 0; v < 8; v++) {
        const double g[3] = {hex8_cx(v) ? 0.25 : -0.25, hex8_cy(v) ? 0.25 : -0.25, hex8_cz(v) ? 0.25 : -0.25};
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < 3; j++) J[d * 3 + j] += c[d][v] * g[j];
    }
}

void adjugate3(const double *J, double *A, double *det) {
#pragma clang fp contract(off)
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

void fff_of(const double *J, double scale, float *out) {
#pragma clang fp contract(off)
    double A[9], det;
    adjugate3(J, A, &det);
    int k = 0;
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
            double s = 0;
            for (int c = 0; c < 3; c++) s += A[a * 3 + c] * A[b * 3 + c];
            out[k++] = (float)(s / det * scale);
        }
}

// rows of the 8 hex8 corners in a connectivity of the given type/level
void corner_rows(int type, int level, int *rows) {
    for (int v = 0; v < 8; v++) {
        if (type == SFEM_AMD_SSHEX8)
            rows[v] = lidx(level, hex8_cx(v) * level, hex8_cy(v) * level, hex8_cz(v) * level);
        else
            rows[v] = v;
    }
}

}  // namespace

extern "C" {

ptrdiff_t sfem_amd_box_nnodes(int nx, int ny, int nz) { return (ptrdiff_t)nx * ny * nz; }

ptrdiff_t sfem_amd_box_nelements(int nx, int ny, int nz, int type) {
    const ptrdiff_t c = (ptrdiff_t)(nx - 1) * (ny - 1) * (nz - 1);
    return type == SFEM_AMD_TET4 ? 6 * c : c;
}

int sfem_amd_box_mesh(int nx, int ny, int nz, double w, double h, double t, int type, int32_t *elements,
                      float *points) {
    SFEM_AMD_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "box mesh needs at least 2 nodes per axis");
    SFEM_AMD_REQUIRE(type == SFEM_AMD_TET4 || type == SFEM_AMD_HEX8, "box mesh: TET4 or HEX8");
    const ptrdiff_t nn = sfem_amd_box_nnodes(nx, ny, nz);
    SFEM_AMD_REQUIRE(nn < INT32_MAX, "box mesh too large for int32 ids");
    const ptrdiff_t ld0 = nz, ld1 = (ptrdiff_t)nz * nx;  // leading_dim = [nz, nz*nx, 1]
    for (int yi = 0; yi < ny; yi++)
        for (int xi = 0; xi < nx; xi++)
            for (int zi = 0; zi < nz; zi++) {
                const ptrdiff_t id = yi * ld1 + xi * ld0 + zi;
                points[id] = linspace_f32(xi, nx, w);
                points[nn + id] = linspace_f32(yi, ny, h);
                points[2 * nn + id] = linspace_f32(zi, nz, t);
            }
    const ptrdiff_t ne = sfem_amd_box_nelements(nx, ny, nz, type);
    static const int tets[6][4] = {{0, 1, 3, 7}, {0, 1, 7, 5}, {0, 4, 5, 7}, {1, 2, 3, 6}, {1, 3, 7, 6}, {1, 5, 6, 7}};
    ptrdiff_t count = 0;
    for (int zi = 0; zi < nz - 1; zi++)
        for (int yi = 0; yi < ny - 1; yi++)
            for (int xi = 0; xi < nx - 1; xi++) {
                int32_t hx[8];
                for (int v = 0; v < 8; v++)
                    hx[v] = (int32_t)((xi + hex8_cx(v)) * ld0 + (yi + hex8_cy(v)) * ld1 + (zi + hex8_cz(v)));
                if (type == SFEM_AMD_HEX8) {
                    for (int v = 0; v < 8; v++) elements[v * ne + count] = hx[v];
                    count++;
                } else {
                    for (int k = 0; k < 6; k++, count++)
                        for (int v = 0; v < 4; v++) elements[v * ne + count] = hx[tets[k][v]];
                }
            }
    return SFEM_SUCCESS;
}

// Hex8 -> sshex8 refinement (the role of smesh::to_semistructured, ssmg/tests/sfem_SSGMGTest.cpp:77-88;
// numbering of the in-tree sshex8_generate_elements + sshex8_hierarchical_renumbering,
// mesh/sfem_hex8_mesh_graph.cpp:448-735,994-1060):
//   raw ids: corners keep the macro ids; edge-interior nodes base + edge*(L-1) + step from the
//   smaller corner id, edges ranked by their (smaller, larger) corner ids (the CRS (i<j) order of
//   the edge graph); side-interior nodes per side ranked by first occurrence (element, side) in the
//   HEX8 side order, walked from the side's smallest-id corner o, u fastest (u, v: o's neighbours on
//   the side, u the one with the smaller id); volume nodes base + e*(L-1)^3 + local, x fastest.
//   hierarchical: corners keep their ids; the other nodes are numbered by first encounter over
//   (level, element, z, y, x) of the level list 1, ..., L/2, L.
ptrdiff_t sfem_amd_sshex8_refine_ex(int L, int hierarchical, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me,
                                    const float *mp, int32_t *elements, float *points) {
    if (L < 1) {
        set_last_error("sshex8 refine: level must be >= 1");
        return -1;
    }
    const int Lm1 = L - 1, nxe = nxe_of(L);

    std::vector<uint64_t> ekeys(12 * (size_t)ne);
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int k = 0; k < 12; k++) {
            const uint32_t a = (uint32_t)me[kEdges[k][0] * ne + e], b = (uint32_t)me[kEdges[k][1] * ne + e];
            ekeys[e * 12 + k] = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
        }
    int64_t nedges = 0;
    const std::vector<int64_t> edge_of = rank_keys(ekeys, /*by_key=*/true, nedges);
    ekeys.clear();

    std::vector<std::array<int32_t, 4>> fkeys(6 * (size_t)ne);
    for (ptrdiff_t e = 0; e < ne; e++)
        for (int f = 0; f < 6; f++) {
            std::array<int32_t, 4> k;
            for (int i = 0; i < 4; i++) k[i] = me[hex8_side(f, i) * ne + e];
            std::sort(k.begin(), k.end());
            fkeys[e * 6 + f] = k;
        }
    int64_t nfaces = 0;
    const std::vector<int64_t> face_of = rank_keys(fkeys, /*by_key=*/false, nfaces);
    fkeys.clear();

    const int64_t base_edge = nnodes;
    const int64_t base_face = base_edge + nedges * Lm1;
    const int64_t base_vol = base_face + nfaces * (int64_t)Lm1 * Lm1;
    const int64_t nraw = base_vol + (int64_t)ne * Lm1 * Lm1 * Lm1;
    if (nraw >= INT32_MAX) {
        set_last_error("sshex8 refine: node count exceeds int32");
        return -1;
    }

    std::vector<int> vx(nxe), vy(nxe), vz(nxe);
    for (int z = 0; z <= L; z++)
        for (int y = 0; y <= L; y++)
            for (int x = 0; x <= L; x++) {
                const int v = lidx(L, x, y, z);
                vx[v] = x;
                vy[v] = y;
                vz[v] = z;
            }

    std::vector<int32_t> pre((size_t)nxe * ne);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t e = 0; e < ne; e++) {
        int32_t c[8];
        for (int v = 0; v < 8; v++) c[v] = me[v * ne + e];
        for (int v = 0; v < nxe; v++) {
            const int x = vx[v], y = vy[v], z = vz[v];
            const int bx = (x == 0 || x == L), by = (y == 0 || y == L), bz = (z == 0 || z == L);
            int64_t id;
            switch (bx + by + bz) {
                case 3:
                    id = c[hex8_corner(x / L, y / L, z / L)];
                    break;
                case 2: {
                    // edge along the non-boundary axis; t from the corner with coordinate 0
                    int a, b, t;
                    if (!bx) {
                        a = hex8_corner(0, y / L, z / L), b = hex8_corner(1, y / L, z / L), t = x;
                    } else if (!by) {
                        a = hex8_corner(x / L, 0, z / L), b = hex8_corner(x / L, 1, z / L), t = y;
                    } else {
                        a = hex8_corner(x / L, y / L, 0), b = hex8_corner(x / L, y / L, 1), t = z;
                    }
                    int k = 0;
                    for (; k < 12; k++)
                        if ((kEdges[k][0] == a && kEdges[k][1] == b) || (kEdges[k][0] == b && kEdges[k][1] == a)) break;
                    const int tt = c[a] < c[b] ? t : L - t;
                    id = base_edge + edge_of[e * 12 + k] * Lm1 + (tt - 1);
                    break;
                }
                case 1: {
                    const int f = hex8_side_of(L, x, y, z);
                    int am = 0;
                    for (int i = 1; i < 4; i++)
                        if (c[hex8_side(f, i)] < c[hex8_side(f, am)]) am = i;
                    const int o = hex8_side(f, am);
                    int nu = hex8_side(f, (am + 1) % 4), nv = hex8_side(f, (am + 3) % 4);
                    if (c[nu] > c[nv]) std::swap(nu, nv);
                    const int oc[3] = {hex8_cx(o) * L, hex8_cy(o) * L, hex8_cz(o) * L};
                    const int uc[3] = {hex8_cx(nu) * L, hex8_cy(nu) * L, hex8_cz(nu) * L};
                    const int vc[3] = {hex8_cx(nv) * L, hex8_cy(nv) * L, hex8_cz(nv) * L};
                    const int p[3] = {x, y, z};
                    int iu = 0, iv = 0;
                    for (int d = 0; d < 3; d++) {
                        if (uc[d] != oc[d]) iu = uc[d] > oc[d] ? p[d] - oc[d] : oc[d] - p[d];
                        if (vc[d] != oc[d]) iv = vc[d] > oc[d] ? p[d] - oc[d] : oc[d] - p[d];
                    }
                    id = base_face + face_of[e * 6 + f] * (int64_t)Lm1 * Lm1 + (int64_t)(iv - 1) * Lm1 + (iu - 1);
                    break;
                }
                default:
                    id = base_vol + (int64_t)e * Lm1 * Lm1 * Lm1 + ((int64_t)(z - 1) * Lm1 + (y - 1)) * Lm1 + (x - 1);
            }
            pre[(size_t)v * ne + e] = (int32_t)id;
        }
    }

    int64_t ntotal = nraw;
    std::vector<int32_t> map;  // raw id -> hierarchical id
    if (hierarchical) {
        int levels[32];
        const int nlev = hierarchical_levels(L, levels);
        // a node's first encounter is in the lowest element holding it, at the coarsest level whose
        // grid holds it; the element loop below runs levels, then elements, then z, y, x
        std::vector<int32_t> first((size_t)nraw, INT32_MAX);
        for (ptrdiff_t e = ne - 1; e >= 0; e--)
            for (int v = 0; v < nxe; v++) first[pre[(size_t)v * ne + e]] = (int32_t)e;
        map.assign((size_t)nraw, -1);
        int64_t next = 0;
        for (int cv = 0; cv < 8; cv++)
            for (ptrdiff_t e = 0; e < ne; e++) {
                const int32_t id = me[cv * ne + e];
                map[id] = id;
                next = std::max<int64_t>(next, id);
            }
        next++;
        for (int k = 1; k < nlev; k++) {
            const int l = levels[k], s = L / l;
            std::vector<int> loc;  // local ids of the level-k nodes (not on a coarser grid), (z, y, x) order
            for (int z = 0; z <= l; z++)
                for (int y = 0; y <= l; y++)
                    for (int x = 0; x <= l; x++)
                        if (hierarchical_level_index(L, levels, nlev, x * s, y * s, z * s) == k)
                            loc.push_back(lidx(L, x * s, y * s, z * s));
            for (ptrdiff_t e = 0; e < ne; e++)
                for (const int v : loc) {
                    const int32_t id = pre[(size_t)v * ne + e];
                    if (first[id] == (int32_t)e && map[id] < 0) map[id] = (int32_t)next++;
                }
        }
        ntotal = next;
    }
    if (!elements) return (ptrdiff_t)ntotal;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 0; i < (ptrdiff_t)nxe * ne; i++) elements[i] = hierarchical ? map[pre[i]] : pre[i];

    if (points) {
        // each node written by the lowest element holding it (deterministic; as the device path)
        std::vector<int32_t> owner((size_t)ntotal, INT32_MAX);
        for (ptrdiff_t e = ne - 1; e >= 0; e--)
            for (int v = 0; v < nxe; v++) owner[elements[(size_t)v * ne + e]] = (int32_t)e;
#pragma omp parallel for schedule(static)
        for (ptrdiff_t e = 0; e < ne; e++) {
#pragma clang fp contract(off)
            double cx[8], cy[8], cz[8];
            for (int v = 0; v < 8; v++) {
                const int32_t n = me[v * ne + e];
                cx[v] = mp[n];
                cy[v] = mp[nnodes + n];
                cz[v] = mp[2 * nnodes + n];
            }
            for (int v = 0; v < nxe; v++) {
                const double qx = (double)vx[v] / L, qy = (double)vy[v] / L, qz = (double)vz[v] / L;
                double px = 0, py = 0, pz = 0;
                for (int k = 0; k < 8; k++) {
                    const double f = (hex8_cx(k) ? qx : 1 - qx) * (hex8_cy(k) ? qy : 1 - qy) * (hex8_cz(k) ? qz : 1 - qz);
                    px += f * cx[k];
                    py += f * cy[k];
                    pz += f * cz[k];
                }
                const int32_t n = elements[(size_t)v * ne + e];
                if (owner[n] != (int32_t)e) continue;
                points[n] = (float)px;
                points[ntotal + n] = (float)py;
                points[2 * ntotal + n] = (float)pz;
            }
        }
    }
    return (ptrdiff_t)ntotal;
}

ptrdiff_t sfem_amd_sshex8_refine(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me, const float *mp,
                                 int32_t *elements, float *points) {
    return sfem_amd_sshex8_refine_ex(L, 1, ne, nnodes, me, mp, elements, points);
}

int sfem_amd_fff(int type, int level, ptrdiff_t ne, ptrdiff_t stride, const int32_t *el, ptrdiff_t nn,
                 const float *p, float *fff) {
    SFEM_AMD_REQUIRE(type == SFEM_AMD_TET4 || type == SFEM_AMD_HEX8 || type == SFEM_AMD_SSHEX8, "fff: element type");
    int rows[8];
    corner_rows(type, level, rows);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t e = 0; e < ne; e++) {
        double J[9];
        if (type == SFEM_AMD_TET4) {
            float px[4], py[4], pz[4];
            for (int i = 0; i < 4; i++) {
                const int32_t v = el[i * stride + e];
                px[i] = p[v];
                py[i] = p[nn + v];
                pz[i] = p[2 * nn + v];
            }
            tet4_fff_f32(px, py, pz, &fff[e * 6]);
        } else {
            double x[8], y[8], z[8];
            for (int v = 0; v < 8; v++) {
                const int32_t n = el[rows[v] * stride + e];
                x[v] = p[n];
                y[v] = p[nn + n];
                z[v] = p[2 * nn + n];
            }
            center_jacobian(x, y, z, J);
            fff_of(J, 1.0, &fff[e * 6]);
        }
    }
    return SFEM_SUCCESS;
}

int sfem_amd_adjugate_det(int type, int level, ptrdiff_t ne, ptrdiff_t stride, const int32_t *el, ptrdiff_t nn,
                          const float *p, float *adj, float *det) {
    SFEM_AMD_REQUIRE(type == SFEM_AMD_HEX8 || type == SFEM_AMD_SSHEX8, "adjugate: HEX8 or SSHEX8");
    int rows[8];
    corner_rows(type, level, rows);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t e = 0; e < ne; e++) {
        double x[8], y[8], z[8], J[9], A[9], d;
        for (int v = 0; v < 8; v++) {
            const int32_t n = el[rows[v] * stride + e];
            x[v] = p[n];
            y[v] = p[nn + n];
            z[v] = p[2 * nn + n];
        }
        center_jacobian(x, y, z, J);
        adjugate3(J, A, &d);
        for (int i = 0; i < 9; i++) adj[e * 9 + i] = (float)A[i];
        det[e] = (float)d;
    }
    return SFEM_SUCCESS;
}

}  // extern "C"
