// Mesh and geometry setup on the device: hex8 -> sshex8 refinement with the hierarchical
// numbering of the host path (mesh.cpp, identical ids), the refined points, and the FFF /
// adjugate caches of a device mesh.  Integer work: radix sorts (hipCUB) rank the unique edges (by
// their corner ids) and sides (by first occurrence), one thread per (element, local node) assigns
// the raw ids, and the hierarchical renumbering sorts the nodes by their first encounter
// (level, element, local node), found with 64-bit atomic minima.  Setup only, deterministic.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "mesh_device.hpp"

namespace sfem_amd {
namespace {

// the 12 hex8 edges as corner pairs (as mesh.cpp)
__device__ __forceinline__ int edge_corner(int k, int end) {
    constexpr int t[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6}, {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    return t[k][end];
}

__global__ void edge_keys_kernel(const ptrdiff_t ne, const int32_t *__restrict__ me, uint64_t *__restrict__ key,
                                 int32_t *__restrict__ pos) {
    const ptrdiff_t n = ne * 12;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / 12;
        const int k = (int)(t - e * 12);
        const uint32_t a = (uint32_t)me[edge_corner(k, 0) * ne + e], b = (uint32_t)me[edge_corner(k, 1) * ne + e];
        key[t] = ((uint64_t)min(a, b) << 32) | max(a, b);
        pos[t] = (int32_t)t;
    }
}

// the sorted four corner ids of face f as (ids 0,1) and (ids 2,3)
__global__ void face_keys_kernel(const ptrdiff_t ne, const int32_t *__restrict__ me, uint64_t *__restrict__ hi,
                                 uint64_t *__restrict__ lo, int32_t *__restrict__ pos) {
    const ptrdiff_t n = ne * 6;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / 6;
        const int f = (int)(t - e * 6);
        uint32_t k[4];
        for (int i = 0; i < 4; i++) k[i] = (uint32_t)me[hex8_side(f, i) * ne + e];
        for (int i = 1; i < 4; i++)
            for (int j = i; j > 0 && k[j - 1] > k[j]; j--) {
                const uint32_t s = k[j];
                k[j] = k[j - 1];
                k[j - 1] = s;
            }
        hi[t] = ((uint64_t)k[0] << 32) | k[1];
        lo[t] = ((uint64_t)k[2] << 32) | k[3];
        pos[t] = (int32_t)t;
    }
}

__global__ void fill_iota_kernel(const ptrdiff_t n, int32_t *__restrict__ a) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        a[i] = (int32_t)i;
}

__global__ void owner_min_kernel(const ptrdiff_t ne, const int nxe, const int32_t *__restrict__ idx,
                                 int32_t *__restrict__ owner) {
    const ptrdiff_t n = ne * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        atomicMin(&owner[idx[t]], (int32_t)(t / nxe));
}

__global__ void gather_u64_kernel(const ptrdiff_t n, const uint64_t *__restrict__ src, const int32_t *__restrict__ at,
                                  uint64_t *__restrict__ dst) {
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        dst[t] = src[at[t]];
}

// sorted position i starts a run of equal keys: head[i] = i (else 0); first = 1 at run starts, stored
// at the occurrence (pos[i]) or, by_key, at the sorted position i
__global__ void run_heads_kernel(const ptrdiff_t n, const uint64_t *__restrict__ k1, const uint64_t *__restrict__ k2,
                                 const int32_t *__restrict__ pos, const int by_key, int32_t *__restrict__ head,
                                 int32_t *__restrict__ first) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const bool h = i == 0 || k1[i] != k1[i - 1] || (k2 && k2[i] != k2[i - 1]);
        head[i] = h ? (int32_t)i : 0;
        first[by_key ? i : pos[i]] = h ? 1 : 0;
    }
}

// occurrence pos[i] -> rank of its key: the rank at its run's first entry (by occurrence of that
// entry, or by_key by its sorted position)
__global__ void run_rank_kernel(const ptrdiff_t n, const int32_t *__restrict__ pos, const int32_t *__restrict__ head,
                                const int by_key, const int32_t *__restrict__ rank, int32_t *__restrict__ out) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        out[pos[i]] = rank[by_key ? head[i] : pos[head[i]]];
}

struct MaxOp {
    __host__ __device__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};

// per (element, local node): the raw id (host path: the OpenMP loop of sfem_amd_sshex8_refine_ex)
__global__ void refine_ids_kernel(const ptrdiff_t ne, const int L, const int32_t *__restrict__ me,
                                  const int32_t *__restrict__ edge_of, const int32_t *__restrict__ face_of,
                                  const int64_t base_edge, const int64_t base_face, const int64_t base_vol,
                                  int32_t *__restrict__ pre) {
    const int Lm1 = L - 1, N1 = L + 1, nxe = N1 * N1 * N1;
    const ptrdiff_t n = ne * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / nxe;
        const int v = (int)(t - e * nxe);
        const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
        int32_t c[8];
        for (int k = 0; k < 8; k++) c[k] = me[k * ne + e];
        const int bx = (x == 0 || x == L), by = (y == 0 || y == L), bz = (z == 0 || z == L);
        int64_t id;
        switch (bx + by + bz) {
            case 3:
                id = c[hex8_corner(x / L, y / L, z / L)];
                break;
            case 2: {
                int a, b, tt;
                if (!bx) {
                    a = hex8_corner(0, y / L, z / L), b = hex8_corner(1, y / L, z / L), tt = x;
                } else if (!by) {
                    a = hex8_corner(x / L, 0, z / L), b = hex8_corner(x / L, 1, z / L), tt = y;
                } else {
                    a = hex8_corner(x / L, y / L, 0), b = hex8_corner(x / L, y / L, 1), tt = z;
                }
                int k = 0;
                for (; k < 12; k++)
                    if ((edge_corner(k, 0) == a && edge_corner(k, 1) == b) ||
                        (edge_corner(k, 0) == b && edge_corner(k, 1) == a))
                        break;
                const int s = c[a] < c[b] ? tt : L - tt;
                id = base_edge + (int64_t)edge_of[e * 12 + k] * Lm1 + (s - 1);
                break;
            }
            case 1: {
                const int f = hex8_side_of(L, x, y, z);
                int am = 0;
                for (int i = 1; i < 4; i++)
                    if (c[hex8_side(f, i)] < c[hex8_side(f, am)]) am = i;
                const int o = hex8_side(f, am);
                int nu = hex8_side(f, (am + 1) % 4), nv = hex8_side(f, (am + 3) % 4);
                if (c[nu] > c[nv]) {
                    const int s = nu;
                    nu = nv;
                    nv = s;
                }
                const int oc[3] = {hex8_cx(o) * L, hex8_cy(o) * L, hex8_cz(o) * L};
                const int uc[3] = {hex8_cx(nu) * L, hex8_cy(nu) * L, hex8_cz(nu) * L};
                const int vc[3] = {hex8_cx(nv) * L, hex8_cy(nv) * L, hex8_cz(nv) * L};
                const int p[3] = {x, y, z};
                int iu = 0, iv = 0;
                for (int d = 0; d < 3; d++) {
                    if (uc[d] != oc[d]) iu = uc[d] > oc[d] ? p[d] - oc[d] : oc[d] - p[d];
                    if (vc[d] != oc[d]) iv = vc[d] > oc[d] ? p[d] - oc[d] : oc[d] - p[d];
                }
                id = base_face + (int64_t)face_of[e * 6 + f] * Lm1 * Lm1 + (int64_t)(iv - 1) * Lm1 + (iu - 1);
                break;
            }
            default:
                id = base_vol + (int64_t)e * Lm1 * Lm1 * Lm1 + ((int64_t)(z - 1) * Lm1 + (y - 1)) * Lm1 + (x - 1);
        }
        pre[t] = (int32_t)id;
    }
}

// First encounter of each non-corner node in the hierarchical renumbering's traversal (levels
// coarse to fine, then elements, then the level's local nodes in (z, y, x) order): the minimum of
// key = (level index, element, local id on the level's grid) over the node's occurrences.
__global__ void first_encounter_kernel(const ptrdiff_t ne, const int L, const int32_t *__restrict__ pre,
                                       unsigned long long *__restrict__ key) {
    int levels[32];
    const int nlev = hierarchical_levels(L, levels);
    const int N1 = L + 1, nxe = N1 * N1 * N1;
    const ptrdiff_t n = ne * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / nxe;
        const int v = (int)(t - e * nxe);
        const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
        const int k = hierarchical_level_index(L, levels, nlev, x, y, z);
        if (k == 0) continue;  // macro corner: keeps its id
        const int l = levels[k], st = L / l;
        const unsigned long long kk = ((unsigned long long)k << 58) | ((unsigned long long)e << 24) |
                                      (unsigned long long)lidx(l, x / st, y / st, z / st);
        atomicMin(&key[pre[t]], kk);
    }
}

// raw ids [base, n) with their keys, for the sort; corners [0, base) map to themselves
__global__ void renumber_pairs_kernel(const ptrdiff_t n, const ptrdiff_t base, const unsigned long long *__restrict__ key,
                                      uint64_t *__restrict__ k_out, int32_t *__restrict__ id_out,
                                      int32_t *__restrict__ perm) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        if (i < base) {
            perm[i] = (int32_t)i;
        } else {
            k_out[i - base] = (uint64_t)key[i];
            id_out[i - base] = (int32_t)i;
        }
    }
}

// sorted position i -> new id next + i
__global__ void renumber_place_kernel(const ptrdiff_t n, const int32_t next, const int32_t *__restrict__ sorted_id,
                                      int32_t *__restrict__ perm) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        perm[sorted_id[i]] = next + (int32_t)i;
}

__global__ void fill_u64_kernel(const ptrdiff_t n, const unsigned long long v, unsigned long long *__restrict__ a) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// element-major connectivity of the refined mesh, and the lowest element holding each node
__global__ void refine_connect_kernel(const ptrdiff_t ne, const int nxe, const int32_t *__restrict__ pre,
                                      const int32_t *__restrict__ perm, int32_t *__restrict__ idx,
                                      int32_t *__restrict__ owner) {
    const ptrdiff_t n = ne * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const int32_t g = perm[pre[t]];
        idx[t] = g;
        if (owner) atomicMin(&owner[g], (int32_t)(t / nxe));
    }
}

// points of the refined nodes: trilinear map of the macro element, written by the node's lowest
// element (the host path writes from every element)
__global__ void refine_points_kernel(const ptrdiff_t ne, const int L, const ptrdiff_t nn_macro, const ptrdiff_t nn,
                                     const int32_t *__restrict__ me, const float *__restrict__ mp,
                                     const int32_t *__restrict__ idx, const int32_t *__restrict__ owner,
                                     float *__restrict__ points) {
#pragma clang fp contract(off)
    const int N1 = L + 1, nxe = N1 * N1 * N1;
    const ptrdiff_t n = ne * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / nxe;
        const int v = (int)(t - e * nxe);
        const int32_t g = idx[t];
        if (owner[g] != (int32_t)e) continue;
        const double qx = (double)(v % N1) / L, qy = (double)((v / N1) % N1) / L, qz = (double)(v / (N1 * N1)) / L;
        double px = 0, py = 0, pz = 0;
        for (int k = 0; k < 8; k++) {
            const int32_t c = me[k * ne + e];
            const double f = (hex8_cx(k) ? qx : 1 - qx) * (hex8_cy(k) ? qy : 1 - qy) * (hex8_cz(k) ? qz : 1 - qz);
            px += f * (double)mp[c];
            py += f * (double)mp[nn_macro + c];
            pz += f * (double)mp[2 * nn_macro + c];
        }
        points[g] = (float)px;
        points[nn + g] = (float)py;
        points[2 * nn + g] = (float)pz;
    }
}

__global__ void fill_i32_kernel(const ptrdiff_t n, const int32_t v, int32_t *__restrict__ a) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// Jacobian at the element centre from its 8 corners (host: center_jacobian), FFF / adjugate
__device__ void centre_jacobian(const double *x, const double *y, const double *z, double *J) {
#pragma clang fp contract(off)
    const double *c[3] = {x, y, z};
    for (int i = 0; i < 9; i++) J[i] = 0;
    for (int v = 0; v < 8; v++) {
        const double g[3] = {hex8_cx(v) ? 0.25 : -0.25, hex8_cy(v) ? 0.25 : -0.25, hex8_cz(v) ? 0.25 : -0.25};
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < 3; j++) J[d * 3 + j] += c[d][v] * g[j];
    }
}
__device__ void adjugate_of(const double *J, double *A, double *det) {
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

// kind 0: FFF [E][6]; kind 1: adjugate [E][9] and det [E].  Corner v of element e is
// idx[e * nv + row_v] (sshex8, element-major) or idx[v * E + e] (hex8/tet4, SoA).
__global__ void geometry_kernel(const ptrdiff_t E, const int type, const int L, const int32_t *__restrict__ idx,
                                const ptrdiff_t nn, const float *__restrict__ p, const int kind,
                                float *__restrict__ out, float *__restrict__ det_out) {
    // no contraction into FMA: the words equal the host path's (mesh.cpp)
#pragma clang fp contract(off)
    for (ptrdiff_t e = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (ptrdiff_t)gridDim.x * blockDim.x) {
        double J[9];
        const double scale = 1.0;
        if (type == SFEM_AMD_TET4) {
            // the reference's float32 evaluation (as the host path)
            float px[4], py[4], pz[4];
            for (int i = 0; i < 4; i++) {
                const int32_t v = idx[i * E + e];
                px[i] = p[v];
                py[i] = p[nn + v];
                pz[i] = p[2 * nn + v];
            }
            if (kind == 0) tet4_fff_f32(px, py, pz, &out[e * 6]);
            continue;
        } else {
            double x[8], y[8], z[8];
            for (int v = 0; v < 8; v++) {
                const int32_t n = type == SFEM_AMD_SSHEX8
                                          ? idx[e * nxe_of(L) + lidx(L, hex8_cx(v) * L, hex8_cy(v) * L, hex8_cz(v) * L)]
                                          : idx[v * E + e];
                x[v] = p[n];
                y[v] = p[nn + n];
                z[v] = p[2 * nn + n];
            }
            centre_jacobian(x, y, z, J);
        }
        double A[9], det;
        adjugate_of(J, A, &det);
        if (kind == 0) {
            int k = 0;
            for (int a = 0; a < 3; a++)
                for (int b = a; b < 3; b++) {
                    double s = 0;
                    for (int c = 0; c < 3; c++) s += A[a * 3 + c] * A[b * 3 + c];
                    out[e * 6 + k++] = (float)(s / det * scale);
                }
        } else {
            for (int i = 0; i < 9; i++) out[e * 9 + i] = (float)A[i];
            det_out[e] = (float)det;
        }
    }
}

// ranks of unique keys: occurrence t (= element * slots + slot) -> rank of its key among the unique
// keys ordered by first occurrence, or (by_key) by key (mesh.cpp: rank_keys).  Keys are (k1, k2)
// (k2 = nullptr: k1 alone); radix sorts are stable, so after sorting equal keys keep occurrence
// order and a run's first entry is the key's first occurrence.
int rank_keys(ptrdiff_t n, const uint64_t *k1, const uint64_t *k2, bool by_key, int32_t *out, int64_t &n_unique,
              hipStream_t s) {
    n_unique = 0;
    if (n == 0) return SFEM_SUCCESS;
    if (n >= INT_MAX) {
        set_last_error("refine on device: too many keys");
        return SFEM_FAILURE;
    }
    const int ni = (int)n;
    DevArray<uint64_t> sk1, sk2, t1;
    DevArray<int32_t> pos, spos, head, headmax, first, rank;
    DevArray<unsigned char> tmp;
    size_t tb = 0;
    auto sort = [&](const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout) -> int {
        size_t need = 0;
        SFEM_AMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, ni, 0, 64, s));
        if (need > tb) {
            if (tmp.alloc(need)) return SFEM_FAILURE;
            tb = need;
        }
        SFEM_AMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kin, kout, vin, vout, ni, 0, 64, s));
        return SFEM_SUCCESS;
    };
    if (pos.alloc(n) || spos.alloc(n) || sk1.alloc(n) || head.alloc(n) || headmax.alloc(n) || first.alloc(n) ||
        rank.alloc(n))
        return SFEM_FAILURE;
    fill_iota_kernel<<<grid_for(n), kBlock, 0, s>>>(n, pos.p);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    if (k2) {
        if (sk2.alloc(n) || t1.alloc(n)) return SFEM_FAILURE;
        if (sort(k2, sk2.p, pos.p, spos.p)) return SFEM_FAILURE;        // by k2
        gather_u64_kernel<<<grid_for(n), kBlock, 0, s>>>(n, k1, spos.p, t1.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        if (sort(t1.p, sk1.p, spos.p, pos.p)) return SFEM_FAILURE;      // then stably by k1
        gather_u64_kernel<<<grid_for(n), kBlock, 0, s>>>(n, k2, pos.p, sk2.p);  // k2 in the final order
        SFEM_AMD_HIP_CHECK(hipGetLastError());
    } else {
        if (sort(k1, sk1.p, pos.p, spos.p)) return SFEM_FAILURE;
        SFEM_AMD_HIP_CHECK(hipMemcpyAsync(pos.p, spos.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    // pos: occurrence of sorted entry i
    run_heads_kernel<<<grid_for(n), kBlock, 0, s>>>(n, sk1.p, k2 ? sk2.p : nullptr, pos.p, by_key ? 1 : 0, head.p,
                                                    first.p);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    size_t need = 0;
    SFEM_AMD_HIP_CHECK(hipcub::DeviceScan::InclusiveScan(nullptr, need, head.p, headmax.p, MaxOp(), ni, s));
    if (need > tb) {
        if (tmp.alloc(need)) return SFEM_FAILURE;
        tb = need;
    }
    SFEM_AMD_HIP_CHECK(hipcub::DeviceScan::InclusiveScan(tmp.p, tb, head.p, headmax.p, MaxOp(), ni, s));
    need = 0;
    SFEM_AMD_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, first.p, rank.p, ni, s));
    if (need > tb) {
        if (tmp.alloc(need)) return SFEM_FAILURE;
        tb = need;
    }
    SFEM_AMD_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, first.p, rank.p, ni, s));
    run_rank_kernel<<<grid_for(n), kBlock, 0, s>>>(n, pos.p, headmax.p, by_key ? 1 : 0, rank.p, out);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    int32_t last[2];
    SFEM_AMD_HIP_CHECK(hipMemcpyAsync(&last[0], rank.p + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SFEM_AMD_HIP_CHECK(hipMemcpyAsync(&last[1], first.p + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    n_unique = (int64_t)last[0] + last[1];
    return SFEM_SUCCESS;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
int sshex8_refine_device(int L, ptrdiff_t ne, ptrdiff_t nnodes, const int32_t *me, DevArray<int32_t> &idx,
                         ptrdiff_t &ntotal_out, hipStream_t s) {
    if (L < 1) {
        set_last_error("sshex8 refine: level must be >= 1");
        return SFEM_FAILURE;
    }
    const int Lm1 = L - 1, nxe = nxe_of(L);
    DevArray<uint64_t> ek, fk1, fk2;
    DevArray<int32_t> edge_of, face_of, dummy;
    int64_t nedges = 0, nfaces = 0;
    if (ne > 0) {
        if (ek.alloc((size_t)12 * ne) || edge_of.alloc((size_t)12 * ne) || dummy.alloc((size_t)12 * ne))
            return SFEM_FAILURE;
        edge_keys_kernel<<<grid_for(12 * ne), kBlock, 0, s>>>(ne, me, ek.p, dummy.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        if (rank_keys(12 * ne, ek.p, nullptr, /*by_key=*/true, edge_of.p, nedges, s)) return SFEM_FAILURE;
        ek.reset();
        if (fk1.alloc((size_t)6 * ne) || fk2.alloc((size_t)6 * ne) || face_of.alloc((size_t)6 * ne))
            return SFEM_FAILURE;
        face_keys_kernel<<<grid_for(6 * ne), kBlock, 0, s>>>(ne, me, fk1.p, fk2.p, dummy.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        if (rank_keys(6 * ne, fk1.p, fk2.p, /*by_key=*/false, face_of.p, nfaces, s)) return SFEM_FAILURE;
    }
    const int64_t base_edge = nnodes;
    const int64_t base_face = base_edge + nedges * Lm1;
    const int64_t base_vol = base_face + nfaces * (int64_t)Lm1 * Lm1;
    const int64_t nraw = base_vol + (int64_t)ne * Lm1 * Lm1 * Lm1;
    if (nraw >= INT32_MAX || (int64_t)ne * nxe >= INT32_MAX || ne >= ((ptrdiff_t)1 << 34)) {
        set_last_error("sshex8 refine: node count exceeds int32");
        return SFEM_FAILURE;
    }
    DevArray<int32_t> pre, perm;
    if (pre.alloc((size_t)nxe * ne) || perm.alloc((size_t)nraw) || idx.alloc((size_t)nxe * ne)) return SFEM_FAILURE;
    if (ne > 0) {
        refine_ids_kernel<<<grid_for((ptrdiff_t)nxe * ne), kBlock, 0, s>>>(ne, L, me, edge_of.p, face_of.p, base_edge,
                                                                           base_face, base_vol, pre.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
    }
    // hierarchical renumbering: corners (raw ids below nnodes, the macro ids) keep their ids; the other
    // nodes take next, next+1, ... in the order of their first-encounter keys (next = largest corner
    // id + 1, as sshex8_hierarchical_renumbering, mesh/sfem_hex8_mesh_graph.cpp:1006-1025)
    int32_t next = 0;
    if (ne > 0) {
        DevArray<int32_t> mx;
        DevArray<unsigned char> tmp;
        if (mx.alloc(1)) return SFEM_FAILURE;
        size_t need = 0;
        SFEM_AMD_HIP_CHECK(hipcub::DeviceReduce::Max(nullptr, need, me, mx.p, (int)(8 * ne), s));
        if (tmp.alloc(need)) return SFEM_FAILURE;
        SFEM_AMD_HIP_CHECK(hipcub::DeviceReduce::Max(tmp.p, need, me, mx.p, (int)(8 * ne), s));
        SFEM_AMD_HIP_CHECK(hipMemcpyAsync(&next, mx.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
        next += 1;
    }
    const ptrdiff_t nrest = (ptrdiff_t)(nraw - nnodes);
    if (nrest > 0) {
        DevArray<unsigned long long> key;
        DevArray<uint64_t> k_in, k_out;
        DevArray<int32_t> id_in, id_out;
        DevArray<unsigned char> tmp;
        if (key.alloc((size_t)nraw) || k_in.alloc((size_t)nrest) || k_out.alloc((size_t)nrest) ||
            id_in.alloc((size_t)nrest) || id_out.alloc((size_t)nrest))
            return SFEM_FAILURE;
        fill_u64_kernel<<<grid_for(nraw), kBlock, 0, s>>>(nraw, ~0ull, key.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        first_encounter_kernel<<<grid_for((ptrdiff_t)nxe * ne), kBlock, 0, s>>>(ne, L, pre.p, key.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        renumber_pairs_kernel<<<grid_for(nraw), kBlock, 0, s>>>(nraw, nnodes, key.p, k_in.p, id_in.p, perm.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
        size_t need = 0;
        SFEM_AMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, k_in.p, k_out.p, id_in.p, id_out.p,
                                                              (int)nrest, 0, 64, s));
        if (tmp.alloc(need)) return SFEM_FAILURE;
        SFEM_AMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, need, k_in.p, k_out.p, id_in.p, id_out.p,
                                                              (int)nrest, 0, 64, s));
        renumber_place_kernel<<<grid_for(nrest), kBlock, 0, s>>>(nrest, next, id_out.p, perm.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
    } else if (nnodes > 0) {
        renumber_pairs_kernel<<<grid_for(nraw), kBlock, 0, s>>>(nraw, nnodes, nullptr, nullptr, nullptr, perm.p);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
    }
    ntotal_out = (ptrdiff_t)next + nrest;
    if (ne > 0) {
        refine_connect_kernel<<<grid_for((ptrdiff_t)nxe * ne), kBlock, 0, s>>>(ne, nxe, pre.p, perm.p, idx.p, nullptr);
        SFEM_AMD_HIP_CHECK(hipGetLastError());
    }
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    return SFEM_SUCCESS;
}

int sshex8_refine_points_device(int L, ptrdiff_t ne, ptrdiff_t nn_macro, ptrdiff_t nn, const int32_t *me,
                                const float *mp, const int32_t *idx, float *points, hipStream_t s) {
    if (ne == 0) return SFEM_SUCCESS;
    const int nxe = nxe_of(L);
    DevArray<int32_t> owner;
    if (owner.alloc((size_t)nn)) return SFEM_FAILURE;
    fill_i32_kernel<<<grid_for(nn), kBlock, 0, s>>>(nn, INT32_MAX, owner.p);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    owner_min_kernel<<<grid_for((ptrdiff_t)nxe * ne), kBlock, 0, s>>>(ne, nxe, idx, owner.p);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    refine_points_kernel<<<grid_for((ptrdiff_t)nxe * ne), kBlock, 0, s>>>(ne, L, nn_macro, nn, me, mp, idx, owner.p,
                                                                          points);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    return SFEM_SUCCESS;
}

int mesh_geometry_device(int type, int L, ptrdiff_t E, const int32_t *idx, ptrdiff_t nn, const float *p, int kind,
                         float *out, float *det, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    geometry_kernel<<<grid_for(E), kBlock, 0, s>>>(E, type, L, idx, nn, p, kind, out, det);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

}  // namespace sfem_amd

namespace sfem_amd {
namespace {
__global__ void aos_to_soa_kernel(const ptrdiff_t E, const int k, const float *__restrict__ aos, float *__restrict__ soa) {
    const ptrdiff_t n = E * k;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / k;
        const int q = (int)(t - e * k);
        soa[q * E + e] = aos[t];
    }
}
// round to nearest even, as to_bf16 on the host (NaN/Inf keep their upper half)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__global__ void aos_to_bf16_soa_kernel(const ptrdiff_t E, const int k, const float *__restrict__ aos,
                                       const float *__restrict__ coeff, uint16_t *__restrict__ soa) {
    const ptrdiff_t n = E * k;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / k;
        const int q = (int)(t - e * k);
        const double c = coeff ? (double)coeff[e] : 1.0;
        soa[q * E + e] = bf16_rne((float)((double)aos[t] * c));
    }
}
}  // namespace

int aos_to_soa_device(const float *aos, ptrdiff_t E, int k, float *soa, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    aos_to_soa_kernel<<<grid_for(E * k), kBlock, 0, s>>>(E, k, aos, soa);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int aos_to_bf16_soa_device(const float *aos, const float *coeff, ptrdiff_t E, int k, uint16_t *soa, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    aos_to_bf16_soa_kernel<<<grid_for(E * k), kBlock, 0, s>>>(E, k, aos, coeff, soa);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
}  // namespace sfem_amd
