// gfx950 kernels: sshex8 transfers, BLAS-1 (deterministic reductions), fused solver
// updates and Dirichlet helpers.  T = double (plain method) or float (FP32 preconditioner).
#include <initializer_list>
#include <mutex>
#include <string>

#include "common.hpp"
#include "transfer_blas_kernels.hpp"

namespace sfem_amd {

// ------------------------------------------------------------------ transfers
// Nested levels share node ids (hierarchical numbering): the node at coarse-local
// (X,Y,Z) of level Lc is the fine-local node (X*s, Y*s, Z*s) of the fine level Lf = s*Lc,
// so both directions read the fine element-major connectivity only (the reference's
// from_level_stride, operators/sshex8/sshex8_interpolate.cpp:259-632).

// Fine value = trilinear interpolant of the coarse cell corners (the reference's x/y/z
// edge, face and volume passes compose to this).  Every fine node is written by one
// element only (owner = lowest element id containing it), so results are deterministic.
template <typename T>
__global__ void sshex8_prolongate_kernel(const ptrdiff_t E, const int Lf, const int Lc, const int32_t *__restrict__ idx,
                                         const uint8_t *__restrict__ own, const int vs, const T *__restrict__ from,
                                         T *__restrict__ to) {
    const int N1 = Lf + 1, nxe = N1 * N1 * N1, s = Lf / Lc;
    const ptrdiff_t n = E * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        if (!own[t]) continue;
        const ptrdiff_t e = t / nxe;
        const int v = (int)(t - e * nxe);
        const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
        int cx = x / s, cy = y / s, cz = z / s;
        if (cx == Lc) cx--;
        if (cy == Lc) cy--;
        if (cz == Lc) cz--;
        const double tx = (double)(x - cx * s) / s, ty = (double)(y - cy * s) / s, tz = (double)(z - cz * s) / s;
        const int32_t *row = idx + e * nxe;
        const ptrdiff_t g = row[v];
        for (int d = 0; d < vs; d++) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int bx = hex8_cx(k), by = hex8_cy(k), bz = hex8_cz(k);
                const double w = (bx ? tx : 1 - tx) * (by ? ty : 1 - ty) * (bz ? tz : 1 - tz);
                if (w == 0) continue;
                const int32_t cg = row[lidx(Lf, (cx + bx) * s, (cy + by) * s, (cz + bz) * s)];
                acc += w * (double)from[(ptrdiff_t)cg * vs + d];
            }
            to[g * vs + d] = (T)acc;
        }
    }
}

// Coarse += P^T fine: each coarse-local node sums the weighted fine nodes of its
// (2s-1)^3 neighbourhood inside the element, fine values divided by their element
// incidence count as sshex8_restrict does (operators/sshex8/sshex8_interpolate.cpp:633-767),
// then one atomic per (element, coarse node).
template <typename T>
__global__ void sshex8_restrict_kernel(const ptrdiff_t E, const int Lf, const int Lc, const int32_t *__restrict__ idx,
                                       const uint16_t *__restrict__ count, const int vs, const T *__restrict__ from,
                                       T *__restrict__ to) {
    const int N1 = Lf + 1, nxe = N1 * N1 * N1, s = Lf / Lc, C1 = Lc + 1, cnxe = C1 * C1 * C1;
    const ptrdiff_t n = E * cnxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / cnxe;
        const int c = (int)(t - e * cnxe);
        const int X = c % C1, Y = (c / C1) % C1, Z = c / (C1 * C1);
        const int32_t *row = idx + e * nxe;
        const int x0 = X * s - (s - 1), y0 = Y * s - (s - 1), z0 = Z * s - (s - 1);
        for (int d = 0; d < vs; d++) {
            double acc = 0;
            for (int k = 0; k < 2 * s - 1; k++) {
                const int z = z0 + k;
                if (z < 0 || z > Lf) continue;
                const double wz = 1.0 - (double)abs(z - Z * s) / s;
                for (int j = 0; j < 2 * s - 1; j++) {
                    const int y = y0 + j;
                    if (y < 0 || y > Lf) continue;
                    const double wy = 1.0 - (double)abs(y - Y * s) / s;
                    for (int i = 0; i < 2 * s - 1; i++) {
                        const int x = x0 + i;
                        if (x < 0 || x > Lf) continue;
                        const double wx = 1.0 - (double)abs(x - X * s) / s;
                        const int32_t g = row[lidx(Lf, x, y, z)];
                        acc += wx * wy * wz * ((double)from[(ptrdiff_t)g * vs + d] / count[g]);
                    }
                }
            }
            const int32_t cg = row[lidx(Lf, X * s, Y * s, Z * s)];
            atomicAdd(&to[(ptrdiff_t)cg * vs + d], (T)acc);
        }
    }
}

// LDS-staged variants, one macro element per block.
// Prolongation: the element's coarse values are staged in LDS once, then every owned fine node
// interpolates from LDS (the per-node variant above re-reads up to 8 coarse values and their ids
// from memory for every fine node).
template <typename T, int VS>
__global__ void __launch_bounds__(256) sshex8_prolongate_lds_kernel(const ptrdiff_t E, const int Lf, const int Lc,
                                                                    const int32_t *__restrict__ idx,
                                                                    const uint8_t *__restrict__ own,
                                                                    const T *__restrict__ from, T *__restrict__ to,
                                                                    const uint8_t *__restrict__ keep, const int add) {
    extern __shared__ double tr_smem[];
    const ptrdiff_t e = blockIdx.x;
    if (e >= E) return;  // uniform per block
    const int N1 = Lf + 1, nxe = N1 * N1 * N1, s = Lf / Lc, C1 = Lc + 1, cnxe = C1 * C1 * C1;
    const int32_t *row = idx + e * nxe;
    // NB nodes per thread at a time (coarse nodes here, fine nodes below): their ids, then their values
    constexpr int NB = 4;
    for (int c0 = threadIdx.x; c0 < cnxe; c0 += NB * blockDim.x) {
        ptrdiff_t cg[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int c = c0 + b * blockDim.x, cc = c < cnxe ? c : 0;
            const int X = cc % C1, Y = (cc / C1) % C1, Z = cc / (C1 * C1);
            cg[b] = row[lidx(Lf, X * s, Y * s, Z * s)];
        }
        T fv[NB][VS];
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int d = 0; d < VS; d++) fv[b][d] = from[cg[b] * VS + d];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int c = c0 + b * blockDim.x;
            if (c < cnxe)
#pragma unroll
                for (int d = 0; d < VS; d++) tr_smem[c * VS + d] = (double)fv[b][d];
        }
    }
    __syncthreads();
    const uint8_t *ow = own + e * nxe;
    // NB fine nodes per thread at a time: their owner flags and ids, then (adding) their keep flags
    // and old values, each batch of loads issued together (unowned nodes read valid entries they
    // do not use)
    for (int v0 = threadIdx.x; v0 < nxe; v0 += NB * blockDim.x) {
        uint8_t o[NB];
        ptrdiff_t gg[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int v = v0 + b * blockDim.x;
            const bool in = v < nxe;
            o[b] = in ? ow[v] : 0;
            gg[b] = row[in ? v : 0];
        }
        unsigned km[NB];
        T old[NB][VS];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            km[b] = (add && keep) ? keep[gg[b]] : 0u;
#pragma unroll
            for (int d = 0; d < VS; d++) old[b][d] = add ? to[gg[b] * VS + d] : T(0);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if (!o[b]) continue;
            const int v = v0 + b * blockDim.x;
            const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
            int cx = x / s, cy = y / s, cz = z / s;
            if (cx == Lc) cx--;
            if (cy == Lc) cy--;
            if (cz == Lc) cz--;
            const double tx = (double)(x - cx * s) / s, ty = (double)(y - cy * s) / s, tz = (double)(z - cz * s) / s;
            double acc[VS];
#pragma unroll
            for (int d = 0; d < VS; d++) acc[d] = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int bx = hex8_cx(k), by = hex8_cy(k), bz = hex8_cz(k);
                const double w = (bx ? tx : 1 - tx) * (by ? ty : 1 - ty) * (bz ? tz : 1 - tz);
                if (w == 0) continue;
                const double *src = tr_smem + lidx(Lc, cx + bx, cy + by, cz + bz) * VS;
#pragma unroll
                for (int d = 0; d < VS; d++) acc[d] += w * src[d];
            }
            const ptrdiff_t g = gg[b];
#pragma unroll
            for (int d = 0; d < VS; d++) {
                if (!add)
                    to[g * VS + d] = (T)acc[d];
                else if (!((km[b] >> d) & 1u))
                    to[g * VS + d] = old[b][d] + (T)acc[d];
            }
        }
    }
}

// Restriction: the element's fine values divided by their incidence counts are staged in LDS;
// each coarse node of the element sums its (2s-1)^3 neighbourhood from LDS.  Coarse nodes inside
// the element (owned by it alone) are updated in place; coarse skeleton nodes store their
// element contribution (scalar: node-sorted position pos[e][slot]; vector: element-major slot)
// for the skeleton gather.  No atomics, deterministic.
template <typename T, int VS>
__global__ void __launch_bounds__(256) sshex8_restrict_lds_kernel(const ptrdiff_t E, const int Lf, const int Lc,
                                                                  const int32_t *__restrict__ idx,
                                                                  const uint16_t *__restrict__ count,
                                                                  const T *__restrict__ from, T *__restrict__ to,
                                                                  T *__restrict__ part,
                                                                  const int32_t *__restrict__ cpos,
                                                                  const int overwrite) {
    extern __shared__ double tr_smem[];
    const ptrdiff_t e = blockIdx.x;
    if (e >= E) return;  // uniform per block
    const int N1 = Lf + 1, nxe = N1 * N1 * N1, s = Lf / Lc, C1 = Lc + 1, cnxe = C1 * C1 * C1;
    const int nskc = sk_count(Lc);
    const int32_t *row = idx + e * nxe;
    // NB fine nodes per thread at a time: their ids, then their counts and values together
    constexpr int NB = 4;
    for (int v0 = threadIdx.x; v0 < nxe; v0 += NB * blockDim.x) {
        ptrdiff_t gg[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int v = v0 + b * blockDim.x;
            gg[b] = row[v < nxe ? v : 0];
        }
        double inv[NB], fv[NB][VS];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            inv[b] = 1.0 / (double)count[gg[b]];
#pragma unroll
            for (int d = 0; d < VS; d++) fv[b][d] = (double)from[gg[b] * VS + d];
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int v = v0 + b * blockDim.x;
            if (v < nxe)
#pragma unroll
                for (int d = 0; d < VS; d++) tr_smem[v * VS + d] = fv[b][d] * inv[b];
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < cnxe; c += blockDim.x) {
        const int X = c % C1, Y = (c / C1) % C1, Z = c / (C1 * C1);
        // the node's id and skeleton slot position are loaded before the sums from LDS, which hide
        // their latency
        const int q = sk_slot(Lc, X, Y, Z);
        const ptrdiff_t cg = row[lidx(Lf, X * s, Y * s, Z * s)];
        const ptrdiff_t slot = q < 0 ? 0 : (VS == 1 ? (ptrdiff_t)cpos[e * nskc + q] : (e * nskc + q));
        double acc[VS];
#pragma unroll
        for (int d = 0; d < VS; d++) acc[d] = 0;
        for (int k = -(s - 1); k <= s - 1; k++) {
            const int z = Z * s + k;
            if (z < 0 || z > Lf) continue;
            const double wz = 1.0 - (double)abs(k) / s;
            for (int j = -(s - 1); j <= s - 1; j++) {
                const int y = Y * s + j;
                if (y < 0 || y > Lf) continue;
                const double wy = 1.0 - (double)abs(j) / s;
                for (int i = -(s - 1); i <= s - 1; i++) {
                    const int x = X * s + i;
                    if (x < 0 || x > Lf) continue;
                    const double w = (1.0 - (double)abs(i) / s) * wy * wz;
                    const double *src = tr_smem + lidx(Lf, x, y, z) * VS;
#pragma unroll
                    for (int d = 0; d < VS; d++) acc[d] += w * src[d];
                }
            }
        }
        if (q < 0) {
#pragma unroll
            for (int d = 0; d < VS; d++) to[cg * VS + d] = overwrite ? (T)acc[d] : to[cg * VS + d] + (T)acc[d];
        } else {
#pragma unroll
            for (int d = 0; d < VS; d++) part[slot * VS + d] = (T)acc[d];
        }
    }
}

__global__ void node_owner_kernel(const ptrdiff_t E, const int nxe, const int32_t *__restrict__ idx,
                                  int32_t *__restrict__ owner) {
    const ptrdiff_t n = E * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        atomicMin(&owner[idx[t]], (int32_t)(t / nxe));
}

__global__ void owner_mask_kernel(const ptrdiff_t E, const int nxe, const int32_t *__restrict__ idx,
                                  const int32_t *__restrict__ owner, uint8_t *__restrict__ own) {
    const ptrdiff_t n = E * nxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        own[t] = owner[idx[t]] == (int32_t)(t / nxe) ? 1 : 0;
}

__global__ void incidence_kernel(const ptrdiff_t n, const int32_t *__restrict__ idx, uint32_t *__restrict__ count) {
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        atomicAdd(&count[idx[t]], 1u);
}

__global__ void narrow_u16_kernel(const ptrdiff_t n, const uint32_t *__restrict__ a, uint16_t *__restrict__ b) {
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        b[t] = (uint16_t)a[t];
}

__global__ void fill_i32_kernel(const ptrdiff_t n, int32_t v, int32_t *__restrict__ a) {
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x)
        a[t] = v;
}

// element-major [E][nxe_f] -> rows (X*s,Y*s,Z*s) as element-major [E][nxe_c]
__global__ void derefine_kernel(const ptrdiff_t E, const int Lf, const int Lc, const int32_t *__restrict__ fine,
                                int32_t *__restrict__ coarse) {
    const int N1 = Lf + 1, nxe = N1 * N1 * N1, s = Lf / Lc, C1 = Lc + 1, cnxe = C1 * C1 * C1;
    const ptrdiff_t n = E * cnxe;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t e = t / cnxe;
        const int c = (int)(t - e * cnxe);
        const int X = c % C1, Y = (c / C1) % C1, Z = c / (C1 * C1);
        coarse[t] = fine[e * nxe + lidx(Lf, X * s, Y * s, Z * s)];
    }
}

// element-major sshex8 -> SoA hex8 [8][E] (standard corner order)
__global__ void to_hex8_kernel(const ptrdiff_t E, const int L, const int32_t *__restrict__ ss, int32_t *__restrict__ h8) {
    const int nxe = nxe_of(L);
    const ptrdiff_t n = E * 8;
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const int v = (int)(t / E);
        const ptrdiff_t e = t - (ptrdiff_t)v * E;
        h8[t] = ss[e * nxe + lidx(L, hex8_cx(v) * L, hex8_cy(v) * L, hex8_cz(v) * L)];
    }
}

template <typename T>
int sshex8_prolongate(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint8_t *own, int vs, const T *from,
                      T *to, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    sshex8_prolongate_kernel<T><<<grid_for(E * nxe_of(Lf)), kBlock, 0, s>>>(E, Lf, Lc, idx, own, vs, from, to);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int sshex8_restrict(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint16_t *count, int vs, const T *from,
                    T *to, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    sshex8_restrict_kernel<T><<<grid_for(E * nxe_of(Lc)), kBlock, 0, s>>>(E, Lf, Lc, idx, count, vs, from, to);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

size_t sshex8_transfer_lds_bytes(int L, int vs) { return sizeof(double) * (size_t)nxe_of(L) * vs; }

template <typename T>
int sshex8_prolongate_lds(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint8_t *own, int vs, const T *from,
                          T *to, hipStream_t s, const uint8_t *keep, int add) {
    if (E == 0) return SFEM_SUCCESS;
    const size_t lds = sshex8_transfer_lds_bytes(Lc, vs);
    if (vs == 1)
        sshex8_prolongate_lds_kernel<T, 1><<<(unsigned)E, 256, lds, s>>>(E, Lf, Lc, idx, own, from, to, keep, add);
    else if (vs == 3)
        sshex8_prolongate_lds_kernel<T, 3><<<(unsigned)E, 256, lds, s>>>(E, Lf, Lc, idx, own, from, to, keep, add);
    else if (!add)
        return sshex8_prolongate<T>(E, Lf, Lc, idx, own, vs, from, to, s);
    else {
        set_last_error("sshex8_prolongate_lds: adding form needs vector size 1 or 3");
        return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int sshex8_restrict_lds(ptrdiff_t E, int Lf, int Lc, const int32_t *idx, const uint16_t *count, int vs, const T *from,
                        T *to, T *part, const int32_t *cpos, hipStream_t s, int overwrite) {
    if (E == 0) return SFEM_SUCCESS;
    const size_t lds = sshex8_transfer_lds_bytes(Lf, vs);
    if (vs == 1)
        sshex8_restrict_lds_kernel<T, 1>
                <<<(unsigned)E, 256, lds, s>>>(E, Lf, Lc, idx, count, from, to, part, cpos, overwrite);
    else if (vs == 3)
        sshex8_restrict_lds_kernel<T, 3>
                <<<(unsigned)E, 256, lds, s>>>(E, Lf, Lc, idx, count, from, to, part, cpos, overwrite);
    else {
        set_last_error("sshex8_restrict_lds: vector size 1 or 3");
        return SFEM_FAILURE;
    }
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template int sshex8_prolongate<double>(ptrdiff_t, int, int, const int32_t *, const uint8_t *, int, const double *,
                                       double *, hipStream_t);
template int sshex8_prolongate<float>(ptrdiff_t, int, int, const int32_t *, const uint8_t *, int, const float *,
                                      float *, hipStream_t);
template int sshex8_restrict<double>(ptrdiff_t, int, int, const int32_t *, const uint16_t *, int, const double *,
                                     double *, hipStream_t);
template int sshex8_restrict<float>(ptrdiff_t, int, int, const int32_t *, const uint16_t *, int, const float *,
                                    float *, hipStream_t);
template int sshex8_prolongate_lds<double>(ptrdiff_t, int, int, const int32_t *, const uint8_t *, int, const double *,
                                           double *, hipStream_t, const uint8_t *, int);
template int sshex8_prolongate_lds<float>(ptrdiff_t, int, int, const int32_t *, const uint8_t *, int, const float *,
                                          float *, hipStream_t, const uint8_t *, int);
template int sshex8_restrict_lds<double>(ptrdiff_t, int, int, const int32_t *, const uint16_t *, int, const double *,
                                         double *, double *, const int32_t *, hipStream_t, int);
template int sshex8_restrict_lds<float>(ptrdiff_t, int, int, const int32_t *, const uint16_t *, int, const float *,
                                        float *, float *, const int32_t *, hipStream_t, int);

int node_owner_mask(ptrdiff_t E, int nxe, ptrdiff_t nnodes, const int32_t *idx, uint8_t *own, hipStream_t s) {
    int32_t *owner = nullptr;
    SFEM_AMD_HIP_CHECK(hipMallocAsync((void **)&owner, sizeof(int32_t) * (nnodes > 0 ? nnodes : 1), s));
    fill_i32_kernel<<<grid_for(nnodes), kBlock, 0, s>>>(nnodes, INT32_MAX, owner);
    node_owner_kernel<<<grid_for(E * nxe), kBlock, 0, s>>>(E, nxe, idx, owner);
    owner_mask_kernel<<<grid_for(E * nxe), kBlock, 0, s>>>(E, nxe, idx, owner, own);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    SFEM_AMD_HIP_CHECK(hipFreeAsync(owner, s));
    return SFEM_SUCCESS;
}

int node_incidence(ptrdiff_t n_entries, ptrdiff_t nnodes, const int32_t *idx, uint16_t *count, hipStream_t s) {
    uint32_t *c32 = nullptr;
    SFEM_AMD_HIP_CHECK(hipMallocAsync((void **)&c32, sizeof(uint32_t) * (nnodes > 0 ? nnodes : 1), s));
    SFEM_AMD_HIP_CHECK(hipMemsetAsync(c32, 0, sizeof(uint32_t) * nnodes, s));
    incidence_kernel<<<grid_for(n_entries), kBlock, 0, s>>>(n_entries, idx, c32);
    narrow_u16_kernel<<<grid_for(nnodes), kBlock, 0, s>>>(nnodes, c32, count);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    SFEM_AMD_HIP_CHECK(hipFreeAsync(c32, s));
    return SFEM_SUCCESS;
}

int sshex8_derefine(ptrdiff_t E, int Lf, int Lc, const int32_t *fine, int32_t *coarse, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    derefine_kernel<<<grid_for(E * nxe_of(Lc)), kBlock, 0, s>>>(E, Lf, Lc, fine, coarse);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int sshex8_to_hex8(ptrdiff_t E, int L, const int32_t *ss, int32_t *h8, hipStream_t s) {
    if (E == 0) return SFEM_SUCCESS;
    to_hex8_kernel<<<grid_for(E * 8), kBlock, 0, s>>>(E, L, ss, h8);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

// ------------------------------------------------------------------ BLAS-1
// y = a x + b y (x may be FP32 beside an FP64 y: converted exactly, the same arithmetic)
template <typename TX, typename T>
__global__ void axpby_kernel(const ptrdiff_t n, const T a, const TX *__restrict__ x, const T b, T *__restrict__ y) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        y[i] = a * (T)x[i] + b * y[i];
}
template <typename T>
__global__ void zaxpby_kernel(const ptrdiff_t n, const T a, const T *__restrict__ x, const T b, const T *__restrict__ y,
                              T *__restrict__ z) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        z[i] = a * x[i] + b * y[i];
}
template <typename T>
__global__ void scal_kernel(const ptrdiff_t n, const T a, T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        x[i] *= a;
}
template <typename T>
__global__ void values_kernel(const ptrdiff_t n, const T a, T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        x[i] = a;
}
template <typename T>
__global__ void copy_kernel(const ptrdiff_t n, const T *__restrict__ x, T *__restrict__ y) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}
template <typename T>
__global__ void reciprocal_kernel(const ptrdiff_t n, const T a, T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        x[i] = a / x[i];
}
// z = x * y + a * z
template <typename T>
__global__ void xypaz_kernel(const ptrdiff_t n, const T *__restrict__ x, const T *__restrict__ y, const T a,
                             T *__restrict__ z) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        z[i] = x[i] * y[i] + a * z[i];
}
template <typename S, typename D>
__global__ void cast_kernel(const ptrdiff_t n, const S *__restrict__ x, D *__restrict__ y) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        y[i] = (D)x[i];
}
__global__ void cast_accumulate_kernel(const ptrdiff_t n, const float *__restrict__ x, double *__restrict__ y) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        y[i] += (double)x[i];
}
__global__ void cast_accumulate_f32_kernel(const ptrdiff_t n, const double *__restrict__ x, float *__restrict__ y) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        y[i] += (float)x[i];
}
// Chebyshev3 iteration_update (algebra/sfem_Chebyshev3.hpp:37-52): p = beta p - rhs + Ax; x -= alpha p
template <typename T>
__global__ void cheb3_update_kernel(const ptrdiff_t n, const T alpha, const T beta, const T *__restrict__ rhs,
                                    const T *__restrict__ Ax, T *__restrict__ p, T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const T pi = cheb3_direction<T>(p[i], rhs[i], Ax[i], beta, nullptr, i, 0);
        p[i] = pi;
        x[i] = __builtin_fma(-alpha, pi, x[i]);
    }
}

// diagonally scaled variant: p = beta p + dinv (A x - rhs); x -= alpha p
// Chebyshev3 iteration 0: p = (p holds A x ? p - rhs : -rhs), times dinv when given; x -= alpha p
// (x = -alpha p for a zero initial guess, whatever x holds)
template <typename T>
__global__ void cheb3_start_kernel(const ptrdiff_t n, const T alpha, const T *__restrict__ rhs,
                                   const T *__restrict__ dinv, const int have_ax, T *__restrict__ p,
                                   T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        if (have_ax) {
            const T pi = cheb3_direction<T>(T(0), rhs[i], p[i], T(0), dinv, i, 1);
            p[i] = pi;
            x[i] = __builtin_fma(-alpha, pi, x[i]);
        } else {  // zero initial guess: x's contents are not read
            T pi = -rhs[i];
            if (dinv) pi = dinv[i] * pi;
            p[i] = pi;
            x[i] = -(alpha * pi);
        }
    }
}

template <typename T>
__global__ void cheb3_update_diag_kernel(const ptrdiff_t n, const T alpha, const T beta, const T *__restrict__ rhs,
                                         const T *__restrict__ Ax, const T *__restrict__ dinv, T *__restrict__ p,
                                         T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        const T pi = cheb3_direction<T>(p[i], rhs[i], Ax[i], beta, dinv, i, 0);
        p[i] = pi;
        x[i] = __builtin_fma(-alpha, pi, x[i]);
    }
}

// y += M x with a dense row-major FP64 matrix: one wave64 per row, fixed-order shuffle tree
// (coarsest-level direct solve with a precomputed inverse).
template <typename T>
__global__ void __launch_bounds__(256) dense_matvec_kernel(const ptrdiff_t n, const double *__restrict__ M,
                                                           const T *__restrict__ x, T *__restrict__ y) {
    const ptrdiff_t row = (ptrdiff_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;  // uniform per wave
    double v = 0;
    const double *Mr = M + row * n;
    for (ptrdiff_t j = lane; j < n; j += 64) v += Mr[j] * (double)x[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) y[row] += (T)v;
}

// Block-Jacobi (algebra/sfem_ShiftableJacobi.hpp ShiftableBlockSymJacobi::set_diag and
// algebra/openmp/sfem_openmp_ShiftableJacobi.hpp): the symmetric 3x3 blocks [n][6] (xx, xy, xz,
// yy, yz, zz) become full blocks, rows of constrained dofs become unit rows, the blocks are
// inverted and scaled by omega; stored [n][9] row-major.
template <typename T>
__global__ void block_sym_inverse_kernel(const ptrdiff_t n, const T *__restrict__ sym, const uint8_t *__restrict__ mask,
                                         const double omega, T *__restrict__ inv) {
    for (ptrdiff_t b = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += (ptrdiff_t)gridDim.x * blockDim.x) {
        const T *d = sym + b * 6;
        double m[9] = {(double)d[0], (double)d[1], (double)d[2], (double)d[1], (double)d[3],
                       (double)d[4], (double)d[2], (double)d[4], (double)d[5]};
        if (mask) {
            for (int r = 0; r < 3; r++)
                if (mask[b * 3 + r])
                    for (int c = 0; c < 3; c++) m[r * 3 + c] = (r == c) ? 1.0 : 0.0;
        }
        const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
        const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
        const double s = omega / det;
        T *o = inv + b * 9;
        o[0] = (T)(c00 * s);
        o[1] = (T)((m[2] * m[7] - m[1] * m[8]) * s);
        o[2] = (T)((m[1] * m[5] - m[2] * m[4]) * s);
        o[3] = (T)(c01 * s);
        o[4] = (T)((m[0] * m[8] - m[2] * m[6]) * s);
        o[5] = (T)((m[2] * m[3] - m[0] * m[5]) * s);
        o[6] = (T)(c02 * s);
        o[7] = (T)((m[1] * m[6] - m[0] * m[7]) * s);
        o[8] = (T)((m[0] * m[4] - m[1] * m[3]) * s);
    }
}

// y[3b + r] += sum_c inv[b][3r + c] x[3b + c]
template <typename T>
__global__ void block3_apply_kernel(const ptrdiff_t n, const T *__restrict__ inv, const T *__restrict__ x,
                                    T *__restrict__ y) {
    for (ptrdiff_t b = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += (ptrdiff_t)gridDim.x * blockDim.x) {
        const T *m = inv + b * 9;
        const T x0 = x[3 * b], x1 = x[3 * b + 1], x2 = x[3 * b + 2];
        y[3 * b + 0] += m[0] * x0 + m[1] * x1 + m[2] * x2;
        y[3 * b + 1] += m[3] * x0 + m[4] * x1 + m[5] * x2;
        y[3 * b + 2] += m[6] * x0 + m[7] * x1 + m[8] * x2;
    }
}

// Deterministic reductions: a fixed grid (a function of n only) of 256-thread blocks,
// wave64 shuffle trees and one pass over the partials in a fixed order.
constexpr int kRedBlocks = 1024;
inline int red_blocks(ptrdiff_t n) {
    ptrdiff_t b = (n + 4 * kBlock - 1) / (4 * kBlock);
    if (b < 1) b = 1;
    if (b > kRedBlocks) b = kRedBlocks;
    return (int)b;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kBlock / 64; i++) r += part[i];
    }
    return r;
}

template <typename TX, typename TY>
__global__ void __launch_bounds__(kBlock) dot_partial_kernel(const ptrdiff_t n, const TX *__restrict__ x,
                                                             const TY *__restrict__ y, double *__restrict__ part) {
    double acc = 0;
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        acc += (double)x[i] * (double)y[i];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(kBlock) sum_partials_kernel(const int nb, const double *__restrict__ part,
                                                              double *__restrict__ out) {
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) out[0] = r;
}

// CG fused update (algebra/sfem_cg.hpp:26-72 host fusion): x += a p; r -= a Ap; partial (r,r) and,
// with z, partial (r,z) of the updated r at part[kRedBlocks + block]; with rf, the updated r rounded
// to FP32 into rf (the input of an FP32 preconditioner)
template <typename T, typename TZ>
__global__ void __launch_bounds__(kBlock) cg_update_kernel(const ptrdiff_t n, const T a, const T *__restrict__ p,
                                                           const T *__restrict__ Ap, T *__restrict__ x,
                                                           T *__restrict__ r, const TZ *__restrict__ z,
                                                           float *__restrict__ rf, double *__restrict__ part) {
    double acc = 0, acz = 0;
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        const T ri = r[i] - a * Ap[i];
        r[i] = ri;
        if (rf) rf[i] = (float)ri;
        acc += (double)ri * (double)ri;
        if (z) acz += (double)ri * (double)z[i];
    }
    if (part) {
        const double s = block_sum(acc);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
        if (z) {
            __syncthreads();  // block_sum's shared scratch is reused
            const double t = block_sum(acz);
            if (threadIdx.x == 0) part[kRedBlocks + blockIdx.x] = t;
        }
    }
}

// Pinned host words the final reductions write into directly (no copy kernel): host-synchronising
// reductions only, one at a time per process.
static double *pinned_results() {
    static double *h = nullptr;
    if (!h && hipHostMalloc((void **)&h, 8 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        h = nullptr;
    }
    return h;
}

template <typename T>
__global__ void scatter_value_kernel(const ptrdiff_t n, const int32_t *__restrict__ dofs, const T v, T *__restrict__ x) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        x[dofs[i]] = v;
}
template <typename T>
__global__ void copy_dofs_kernel(const ptrdiff_t n, const int32_t *__restrict__ dofs, const T *__restrict__ src,
                                 T *__restrict__ dst) {
    for (ptrdiff_t i = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (ptrdiff_t)gridDim.x * blockDim.x)
        dst[dofs[i]] = src[dofs[i]];
}

// ---- streaming kernels over 16-byte vectors
// When every array of a streaming operation is 16-byte aligned, each thread moves 16 bytes per
// array (4 floats or 2 doubles) with one load or store instruction: fewer, wider memory
// instructions per byte (a one-float-per-thread copy of 171M floats runs at 3.7 TB/s on MI355X,
// a 2-GiB FP64 copy at 6.3 TB/s).  The element arithmetic is the scalar kernels', so the results
// are the same; the last n % W elements go to the first threads of block 0.  Used for FP32 arrays:
// FP64 ones (2 doubles per thread) made the FP64 multigrid solve of 171M unknowns 3.5% slower
// than one double per thread, which already streams at the FP64 copy's rate.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <typename T>
using vec16_t = typename Vec16<T>::type;

template <typename T>
static inline bool aligned16(std::initializer_list<const void *> ps) {
    if (sizeof(T) != 4) return false;
    for (const void *q : ps)
        if (q && ((uintptr_t)q & 15u) != 0) return false;
    return true;
}

template <typename T, typename F>
__global__ void __launch_bounds__(kBlock) stream_vec_kernel(const ptrdiff_t n, const F f) {
    constexpr int W = 16 / sizeof(T);
    const ptrdiff_t nv = n / W;
    for (ptrdiff_t v = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (ptrdiff_t)gridDim.x * blockDim.x)
        f.vec(v);
    const ptrdiff_t t = nv * W + threadIdx.x;
    if (blockIdx.x == 0 && t < n) f.one(t);
}
template <typename T, typename F>
static int launch_stream_vec(ptrdiff_t n, const F &f, hipStream_t s) {
    constexpr int W = 16 / sizeof(T);
    stream_vec_kernel<T, F><<<grid_for(n / W + 1), kBlock, 0, s>>>(n, f);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
#define SFEM_VEC_LOAD(V, ptr, v) (reinterpret_cast<const V *>(ptr)[v])
#define SFEM_VEC_STORE(V, ptr, v, val) (reinterpret_cast<V *>(ptr)[v] = (val))

// y = a x + b y
template <typename T>
struct AxpbyF {
    T a, b;
    const T *x;
    T *y;
    __device__ void one(ptrdiff_t i) const { y[i] = a * x[i] + b * y[i]; }
    __device__ void vec(ptrdiff_t v) const {
        using V = vec16_t<T>;
        const V xv = SFEM_VEC_LOAD(V, x, v), yv = SFEM_VEC_LOAD(V, y, v);
        V o;
#pragma unroll
        for (int k = 0; k < (int)(16 / sizeof(T)); k++) o[k] = a * xv[k] + b * yv[k];
        SFEM_VEC_STORE(V, y, v, o);
    }
};
template <typename T>
struct ValuesF {
    T a;
    T *x;
    __device__ void one(ptrdiff_t i) const { x[i] = a; }
    __device__ void vec(ptrdiff_t v) const {
        using V = vec16_t<T>;
        V o;
#pragma unroll
        for (int k = 0; k < (int)(16 / sizeof(T)); k++) o[k] = a;
        SFEM_VEC_STORE(V, x, v, o);
    }
};
template <typename T>
struct CopyF {
    const T *x;
    T *y;
    __device__ void one(ptrdiff_t i) const { y[i] = x[i]; }
    __device__ void vec(ptrdiff_t v) const {
        using V = vec16_t<T>;
        SFEM_VEC_STORE(V, y, v, SFEM_VEC_LOAD(V, x, v));
    }
};
// Chebyshev3 steps, as cheb3_update_kernel / cheb3_update_diag_kernel (DIAG) and cheb3_start_kernel
template <typename T, bool DIAG>
struct Cheb3UpdateF {
    T alpha, beta;
    const T *rhs, *Ax, *dinv;
    T *p, *x;
    __device__ void one(ptrdiff_t i) const {
        const T pi = cheb3_direction<T>(p[i], rhs[i], Ax[i], beta, DIAG ? dinv : nullptr, i, 0);
        p[i] = pi;
        x[i] = __builtin_fma(-alpha, pi, x[i]);
    }
    __device__ void vec(ptrdiff_t v) const {
        using V = vec16_t<T>;
        const V pv = SFEM_VEC_LOAD(V, p, v), rv = SFEM_VEC_LOAD(V, rhs, v), av = SFEM_VEC_LOAD(V, Ax, v),
                xv = SFEM_VEC_LOAD(V, x, v);
        V dv;
        if (DIAG) dv = SFEM_VEC_LOAD(V, dinv, v);
        V po, xo;
#pragma unroll
        for (int k = 0; k < (int)(16 / sizeof(T)); k++) {
            const T dk = DIAG ? dv[k] : T(0);
            const T pi = cheb3_direction<T>(pv[k], rv[k], av[k], beta, DIAG ? &dk : nullptr, 0, 0);
            po[k] = pi;
            xo[k] = __builtin_fma(-alpha, pi, xv[k]);
        }
        SFEM_VEC_STORE(V, p, v, po);
        SFEM_VEC_STORE(V, x, v, xo);
    }
};
template <typename T>
struct Cheb3StartF {
    T alpha;
    const T *rhs, *dinv;
    int have_ax;
    T *p, *x;
    __device__ void one(ptrdiff_t i) const {
        if (have_ax) {
            const T pi = cheb3_direction<T>(T(0), rhs[i], p[i], T(0), dinv, i, 1);
            p[i] = pi;
            x[i] = __builtin_fma(-alpha, pi, x[i]);
        } else {
            T pi = -rhs[i];
            if (dinv) pi = dinv[i] * pi;
            p[i] = pi;
            x[i] = -(alpha * pi);
        }
    }
    __device__ void vec(ptrdiff_t v) const {
        using V = vec16_t<T>;
        constexpr int W = 16 / sizeof(T);
        const V rv = SFEM_VEC_LOAD(V, rhs, v);
        V dv, pv, xv, po, xo;
        if (dinv) dv = SFEM_VEC_LOAD(V, dinv, v);
        if (have_ax) {
            pv = SFEM_VEC_LOAD(V, p, v);
            xv = SFEM_VEC_LOAD(V, x, v);
#pragma unroll
            for (int k = 0; k < W; k++) {
                const T dk = dinv ? dv[k] : T(0);
                const T pi = cheb3_direction<T>(T(0), rv[k], pv[k], T(0), dinv ? &dk : nullptr, 0, 1);
                po[k] = pi;
                xo[k] = __builtin_fma(-alpha, pi, xv[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < W; k++) {
                T pi = -rv[k];
                if (dinv) pi = dv[k] * pi;
                po[k] = pi;
                xo[k] = -(alpha * pi);
            }
        }
        SFEM_VEC_STORE(V, p, v, po);
        SFEM_VEC_STORE(V, x, v, xo);
    }
};

template <typename T>
int blas_axpby(ptrdiff_t n, T a, const T *x, T b, T *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({x, y})) return launch_stream_vec<T>(n, AxpbyF<T>{a, b, x, y}, s);
    axpby_kernel<T, T><<<grid_for(n), kBlock, 0, s>>>(n, a, x, b, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_zaxpby(ptrdiff_t n, T a, const T *x, T b, const T *y, T *z, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    zaxpby_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, a, x, b, y, z);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_scal(ptrdiff_t n, T a, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    scal_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, a, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_values(ptrdiff_t n, T a, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({x})) return launch_stream_vec<T>(n, ValuesF<T>{a, x}, s);
    values_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, a, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_copy(ptrdiff_t n, const T *x, T *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({x, y})) return launch_stream_vec<T>(n, CopyF<T>{x, y}, s);
    copy_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_reciprocal(ptrdiff_t n, T a, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    reciprocal_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, a, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int blas_xypaz(ptrdiff_t n, const T *x, const T *y, T a, T *z, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    xypaz_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, x, y, a, z);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename S, typename D>
int blas_cast(ptrdiff_t n, const S *x, D *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    cast_kernel<S, D><<<grid_for(n), kBlock, 0, s>>>(n, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int cheb3_update(ptrdiff_t n, T alpha, T beta, const T *rhs, const T *Ax, T *p, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({rhs, Ax, p, x}))
        return launch_stream_vec<T>(n, Cheb3UpdateF<T, false>{alpha, beta, rhs, Ax, nullptr, p, x}, s);
    cheb3_update_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, alpha, beta, rhs, Ax, p, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int cheb3_start(ptrdiff_t n, T alpha, const T *rhs, const T *dinv, int have_ax, T *p, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({rhs, dinv, p, x}))
        return launch_stream_vec<T>(n, Cheb3StartF<T>{alpha, rhs, dinv, have_ax, p, x}, s);
    cheb3_start_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, alpha, rhs, dinv, have_ax, p, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int cheb3_update_diag(ptrdiff_t n, T alpha, T beta, const T *rhs, const T *Ax, const T *dinv, T *p, T *x,
                      hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    if (aligned16<T>({rhs, Ax, dinv, p, x}))
        return launch_stream_vec<T>(n, Cheb3UpdateF<T, true>{alpha, beta, rhs, Ax, dinv, p, x}, s);
    cheb3_update_diag_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, alpha, beta, rhs, Ax, dinv, p, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int block_sym_inverse(ptrdiff_t n, const T *sym, const uint8_t *mask, double omega, T *inv, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    block_sym_inverse_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, sym, mask, omega, inv);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int block3_apply(ptrdiff_t n, const T *inv, const T *x, T *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    block3_apply_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, inv, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int dense_matvec(ptrdiff_t n, const double *M, const T *x, T *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    dense_matvec_kernel<T><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, M, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

// the k sums of partials part[j * kRedBlocks + block] -> host values out[j]
static int finish_sums(int nb, int k, double *ws, double *out, hipStream_t s) {
    double *h = pinned_results();
    for (int j = 0; j < k; j++)
        sum_partials_kernel<<<1, kBlock, 0, s>>>(nb, ws + j * kRedBlocks, h ? h + j : ws + 2 * kRedBlocks + j);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    if (!h)
        SFEM_AMD_HIP_CHECK(hipMemcpyAsync(out, ws + 2 * kRedBlocks, k * sizeof(double), hipMemcpyDeviceToHost, s));
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    if (h)
        for (int j = 0; j < k; j++) out[j] = h[j];
    return SFEM_SUCCESS;
}

template <typename T>
int blas_dot(ptrdiff_t n, const T *x, const T *y, double *ws, double *result_host, hipStream_t s) {
    const int nb = red_blocks(n);
    dot_partial_kernel<T, T><<<nb, kBlock, 0, s>>>(n, x, y, ws);
    return finish_sums(nb, 1, ws, result_host, s);
}

// Allocation-free dot: the per-block partials live in `scratch`, a file-scope device array that
// every GPU holds once for the life of the process (8 KiB: one double per block of the capped
// reduction grid).  Host code never passes the array's name to a kernel: on the host it is a
// stand-in, not a device address.  The address is looked up once per device with that device
// current and kept in scratch_dev[device].  The partials are written in full before they are read,
// so nothing needs clearing; calls that overlap on one GPU share the array and race.
constexpr int kScratchDoubles = kRedBlocks;
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

static double *scratch_ptr() {
    static std::mutex mu;
    static double *scratch_dev[kScratchMaxDevices] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchMaxDevices) {
        (void)hipGetLastError();
        set_last_error("dot_noalloc: no current HIP device (or device index beyond 63)");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p) {
            (void)hipGetLastError();
            set_last_error(std::string("dot_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(err));
            return nullptr;
        }
        scratch_dev[dev] = (double *)p;
    }
    return scratch_dev[dev];
}

template <typename T>
int blas_dot_noalloc(ptrdiff_t n, const T *x, const T *y, double *result, hipStream_t s) {
    if (n < 0 || !result || (n > 0 && (!x || !y))) {
        set_last_error("dot_noalloc: n must be >= 0 and x, y, result non-null");
        return SFEM_FAILURE;
    }
    const int nb = red_blocks(n);  // same grid as blas_dot: the same partials, summed in the same order
    if (nb > kScratchDoubles) {
        set_last_error("dot_noalloc: reduction grid exceeds the 1024-partial scratch array");
        return SFEM_FAILURE;
    }
    double *ws = scratch_ptr();
    if (!ws) return SFEM_FAILURE;
    dot_partial_kernel<T, T><<<nb, kBlock, 0, s>>>(n, x, y, ws);
    sum_partials_kernel<<<1, kBlock, 0, s>>>(nb, ws, result);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
int cg_update(ptrdiff_t n, T a, const T *p, const T *Ap, T *x, T *r, double *ws, double *rtr_host, hipStream_t s,
              const T *z, double *rz_host) {
    const int nb = red_blocks(n);
    cg_update_kernel<T, T><<<nb, kBlock, 0, s>>>(n, a, p, Ap, x, r, rtr_host ? z : nullptr, nullptr,
                                                rtr_host ? ws : nullptr);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    if (!rtr_host) return SFEM_SUCCESS;
    double out[2] = {0, 0};
    if (finish_sums(nb, z ? 2 : 1, ws, out, s)) return SFEM_FAILURE;
    *rtr_host = out[0];
    if (z && rz_host) *rz_host = out[1];
    return SFEM_SUCCESS;
}

// The steps of a flexible CG whose preconditioner runs in FP32 (z stays FP32; the outer vectors
// are FP64): each is the FP64 step with z converted exactly, so the results are those of the
// all-FP64 steps on (double) z.
int blas_dot_mixed(ptrdiff_t n, const double *x, const float *y, double *ws, double *result_host, hipStream_t s) {
    const int nb = red_blocks(n);
    dot_partial_kernel<double, float><<<nb, kBlock, 0, s>>>(n, x, y, ws);
    return finish_sums(nb, 1, ws, result_host, s);
}
int blas_axpby_mixed(ptrdiff_t n, double a, const float *x, double b, double *y, hipStream_t s) {
    axpby_kernel<float, double><<<grid_for(n), kBlock, 0, s>>>(n, a, x, b, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
// x += a p; r -= a Ap; rf = (float) r; *rtr_host = (r, r) and, with zf, *rz_host = (r, zf)
int cg_update_mixed(ptrdiff_t n, double a, const double *p, const double *Ap, double *x, double *r, float *rf,
                    const float *zf, double *ws, double *rtr_host, double *rz_host, hipStream_t s) {
    const int nb = red_blocks(n);
    cg_update_kernel<double, float><<<nb, kBlock, 0, s>>>(n, a, p, Ap, x, r, zf, rf, ws);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    double out[2] = {0, 0};
    if (finish_sums(nb, zf ? 2 : 1, ws, out, s)) return SFEM_FAILURE;
    *rtr_host = out[0];
    if (zf && rz_host) *rz_host = out[1];
    return SFEM_SUCCESS;
}

template <typename T>
int dofs_set_value(ptrdiff_t n, const int32_t *dofs, T v, T *x, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    scatter_value_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, dofs, v, x);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template <typename T>
int dofs_copy(ptrdiff_t n, const int32_t *dofs, const T *src, T *dst, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    copy_dofs_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, dofs, src, dst);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

template <typename T>
__global__ void permute_add_kernel(const ptrdiff_t n, const int bs, const int32_t *__restrict__ perm, const int transpose,
                                   const T *__restrict__ x, T *__restrict__ y) {
    for (ptrdiff_t t = (ptrdiff_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * bs; t += (ptrdiff_t)gridDim.x * blockDim.x) {
        const ptrdiff_t i = t / bs;
        const int c = (int)(t - i * bs);
        const ptrdiff_t j = perm[i];
        if (transpose)
            y[i * bs + c] += x[j * bs + c];
        else
            y[j * bs + c] += x[i * bs + c];
    }
}

template <typename T>
int permute_add(ptrdiff_t n, int bs, const int32_t *perm, int transpose, const T *x, T *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    permute_add_kernel<T><<<grid_for(n * bs), kBlock, 0, s>>>(n, bs, perm, transpose, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}
template int permute_add<double>(ptrdiff_t, int, const int32_t *, int, const double *, double *, hipStream_t);
template int permute_add<float>(ptrdiff_t, int, const int32_t *, int, const float *, float *, hipStream_t);

size_t reduction_workspace_doubles() { return 2 * kRedBlocks + 4; }

int cast_accumulate_f32(ptrdiff_t n, const double *x, float *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    cast_accumulate_f32_kernel<<<grid_for(n), kBlock, 0, s>>>(n, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

int cast_accumulate(ptrdiff_t n, const float *x, double *y, hipStream_t s) {
    if (n <= 0) return SFEM_SUCCESS;
    cast_accumulate_kernel<<<grid_for(n), kBlock, 0, s>>>(n, x, y);
    SFEM_AMD_HIP_CHECK(hipGetLastError());
    return SFEM_SUCCESS;
}

#define INST(T)                                                                                         \
    template int blas_axpby<T>(ptrdiff_t, T, const T *, T, T *, hipStream_t);                           \
    template int blas_zaxpby<T>(ptrdiff_t, T, const T *, T, const T *, T *, hipStream_t);               \
    template int blas_scal<T>(ptrdiff_t, T, T *, hipStream_t);                                          \
    template int blas_values<T>(ptrdiff_t, T, T *, hipStream_t);                                        \
    template int blas_copy<T>(ptrdiff_t, const T *, T *, hipStream_t);                                  \
    template int blas_reciprocal<T>(ptrdiff_t, T, T *, hipStream_t);                                    \
    template int blas_xypaz<T>(ptrdiff_t, const T *, const T *, T, T *, hipStream_t);                   \
    template int cheb3_update<T>(ptrdiff_t, T, T, const T *, const T *, T *, T *, hipStream_t);         \
    template int cheb3_update_diag<T>(ptrdiff_t, T, T, const T *, const T *, const T *, T *, T *, hipStream_t); \
    template int cheb3_start<T>(ptrdiff_t, T, const T *, const T *, int, T *, T *, hipStream_t);          \
    template int dense_matvec<T>(ptrdiff_t, const double *, const T *, T *, hipStream_t);                \
    template int block_sym_inverse<T>(ptrdiff_t, const T *, const uint8_t *, double, T *, hipStream_t);   \
    template int block3_apply<T>(ptrdiff_t, const T *, const T *, T *, hipStream_t);                      \
    template int blas_dot<T>(ptrdiff_t, const T *, const T *, double *, double *, hipStream_t);         \
    template int blas_dot_noalloc<T>(ptrdiff_t, const T *, const T *, double *, hipStream_t);           \
    template int cg_update<T>(ptrdiff_t, T, const T *, const T *, T *, T *, double *, double *, hipStream_t,  \
                              const T *, double *);                                                       \
    template int dofs_set_value<T>(ptrdiff_t, const int32_t *, T, T *, hipStream_t);                    \
    template int dofs_copy<T>(ptrdiff_t, const int32_t *, const T *, T *, hipStream_t);
INST(double)
INST(float)
#undef INST
template int blas_cast<double, float>(ptrdiff_t, const double *, float *, hipStream_t);
template int blas_cast<float, double>(ptrdiff_t, const float *, double *, hipStream_t);

}  // namespace sfem_amd
