// C ABI runtime: device meshes, operators, transfers, BLAS entry points and solver
// handles over the sfem:: templates (include/sfem/).  See include/sfem_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <functional>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "mesh_device.hpp"
#include "sfem/sfem_Chebyshev3.hpp"
#include "sfem/sfem_Multigrid.hpp"
#include "sfem/sfem_Stationary.hpp"
#include "sfem/sfem_cg.hpp"
#include "sshex8_kernels.hpp"
#include "transfer_blas_kernels.hpp"
#include "unstructured_kernels.hpp"

namespace sfem_amd {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

hipStream_t library_stream() {
    constexpr int kMaxDevices = 64;
    static hipStream_t streams[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    if (!streams[dev]) {
        hipStream_t st = nullptr;
        if (hipStreamCreate(&st) != hipSuccess) return nullptr;
        streams[dev] = st;
    }
    return streams[dev];
}

}  // namespace sfem_amd

namespace sfem_amd {

int mesh_counts(const sfem_amd_mesh *m, hipStream_t s) {
    if (m->count.p) return SFEM_SUCCESS;
    if (m->count.alloc(m->nnodes)) return SFEM_FAILURE;
    return node_incidence(m->E * m->nv, m->nnodes, m->idx.p, m->count.p, s);
}
// Skeleton slot lists of an sshex8 mesh viewed at level Lc (Lc divides the mesh level L): the
// coarse-local node (X, Y, Z) of element e is the mesh node at lidx(L, X*s, Y*s, Z*s), s = L/Lc.
// Lc == L is the mesh itself; smaller Lc serve the restriction onto derefined levels.
int mesh_skeleton(const sfem_amd_mesh *m, int Lc) {
    if (Lc <= 0) Lc = m->level;
    auto &slot = m->skel[Lc];
    if (slot) return SFEM_SUCCESS;
    const int L = m->level, nsk = sk_count(Lc), st = L / Lc;
    if (L % Lc) {
        set_last_error("sshex8 skeleton map: level does not divide the mesh level");
        return SFEM_FAILURE;
    }
    const ptrdiff_t E = m->E;
    if ((double)E * nsk >= 2147483647.0) {
        set_last_error("sshex8 skeleton map: E * skeleton nodes per element exceeds int32");
        return SFEM_FAILURE;
    }
    std::vector<int32_t> el((size_t)E * m->nv);
    if (!el.empty())
        SFEM_AMD_HIP_CHECK(hipMemcpy(el.data(), m->idx.p, sizeof(int32_t) * el.size(), hipMemcpyDeviceToHost));
    // mesh-local index of every skeleton slot of the level-Lc view, in slot order
    std::vector<int> sk_lidx(nsk);
    for (int z = 0; z <= Lc; z++)
        for (int y = 0; y <= Lc; y++)
            for (int x = 0; x <= Lc; x++) {
                const int q = sk_slot(Lc, x, y, z);
                if (q >= 0) sk_lidx[q] = lidx(L, x * st, y * st, z * st);
            }
    auto store = std::make_unique<SkeletonStore>();
    if (Lc == L) {
        // the apply kernels update a macro element's interior nodes without atomics: each must
        // belong to that element alone (and to no skeleton); count every node's occurrences
        const int N1 = L + 1;
        std::vector<uint8_t> seen(m->nnodes, 0);
        bool exclusive = true;
        for (ptrdiff_t e = 0; e < E && exclusive; e++)
            for (int v = 0; v < m->nv; v++) {
                const int32_t g = el[(size_t)e * m->nv + v];
                const int x = v % N1, y = (v / N1) % N1, z = v / (N1 * N1);
                if (sk_slot(L, x, y, z) >= 0) {
                    seen[g] |= 1;
                } else {
                    if (seen[g] & 2) exclusive = false;
                    seen[g] |= 2;
                }
            }
        bool covers = exclusive;
        for (ptrdiff_t g = 0; g < m->nnodes && exclusive; g++) {
            if (!seen[g]) covers = false;
            if ((seen[g] & 3) == 3) exclusive = false;  // interior here, skeleton elsewhere
        }
        if (!exclusive) {
            set_last_error("sshex8 mesh: an element-interior node is shared with another element (non-conforming "
                           "connectivity)");
            return SFEM_FAILURE;
        }
        store->covers = covers;
    }
    std::vector<int32_t> cnt(m->nnodes + 1, 0);
    for (ptrdiff_t e = 0; e < E; e++)
        for (int q = 0; q < nsk; q++) cnt[el[(size_t)e * m->nv + sk_lidx[q]]]++;
    std::vector<int32_t> node, off(1, 0), pos(m->nnodes, -1);
    for (ptrdiff_t g = 0; g < m->nnodes; g++)
        if (cnt[g]) {
            pos[g] = (int32_t)node.size();
            node.push_back((int32_t)g);
            off.push_back(off.back() + cnt[g]);
        }
    // slot (e, q) -> its position in the node-sorted list (elements in ascending order), and back
    std::vector<int32_t> fill(node.size(), 0), slot_pos((size_t)E * nsk), ent((size_t)E * nsk);
    for (ptrdiff_t e = 0; e < E; e++)
        for (int q = 0; q < nsk; q++) {
            const int32_t p = pos[el[(size_t)e * m->nv + sk_lidx[q]]];
            const int32_t j = off[p] + fill[p]++;
            slot_pos[(size_t)e * nsk + q] = j;
            ent[j] = (int32_t)(e * nsk + q);
        }
    if (store->node.upload(node.data(), node.size()) || store->off.upload(off.data(), off.size()) ||
        store->pos.upload(slot_pos.data(), slot_pos.size()) || store->ent.upload(ent.data(), ent.size()))
        return SFEM_FAILURE;
    store->n = (ptrdiff_t)node.size();
    slot = std::move(store);
    return SFEM_SUCCESS;
}
// an empty map when the level has none (E == 0: the kernels then return before launching)
static SkeletonMap skeleton_of(const sfem_amd_mesh *m, int Lc = 0) {
    SkeletonMap sk{};
    const auto it = m->skel.find(Lc > 0 ? Lc : m->level);
    if (it == m->skel.end() || !it->second) return sk;
    const SkeletonStore &st = *it->second;
    sk.node = st.node.p;
    sk.off = st.off.p;
    sk.pos = st.pos.p;
    sk.ent = st.ent.p;
    sk.n = st.n;
    return sk;
}
static bool skeleton_covers(const sfem_amd_mesh *m) { return m->skel.at(m->level)->covers; }
static int mesh_owner(const sfem_amd_mesh *m, hipStream_t s) {
    if (m->own.p) return SFEM_SUCCESS;
    if (m->own.alloc((size_t)m->E * m->nv)) return SFEM_FAILURE;
    return node_owner_mask(m->E, m->nv, m->nnodes, m->idx.p, m->own.p, s);
}
static std::atomic<uint64_t> g_state_epoch{0};
uint64_t state_epoch() { return g_state_epoch.load(std::memory_order_acquire); }
void bump_state_epoch() { g_state_epoch.fetch_add(1, std::memory_order_acq_rel); }

bool g_fused_chebyshev = false;
bool g_mixed_cg = true;
bool g_atomic_skeleton = false;
int g_elasticity_form = 0;
bool g_hex8_patches = true;

const sfem_amd_mesh *mesh_subblocks(const sfem_amd_mesh *m, int Ls, hipStream_t s) {
    auto &slot = m->sub[Ls];
    if (slot) return slot.get();
    if (m->type != SFEM_AMD_SSHEX8 || Ls <= 0 || m->level % Ls) {
        set_last_error("sshex8 sub-blocks: the sub-level must divide the mesh level");
        return nullptr;
    }
    const int f = m->level / Ls;
    auto sm = std::make_unique<sfem_amd_mesh>();
    sm->type = SFEM_AMD_SSHEX8;
    sm->level = Ls;
    sm->E = m->E * f * f * f;
    sm->nnodes = m->nnodes;
    sm->nv = nxe_of(Ls);
    if (sm->idx.alloc((size_t)sm->E * sm->nv) || sshex8_subblock_idx(m->E, m->level, Ls, m->idx.p, sm->idx.p, s))
        return nullptr;
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_last_error("sshex8 sub-blocks: stream synchronisation failed");
        return nullptr;
    }
    slot = std::move(sm);
    return slot.get();
}

// Element patches of a tet4/hex8 mesh: breadth-first growth over shared nodes, at most kP elements
// each, the next seed taken from the unassigned neighbours of the previous patch (lowest id first),
// else the lowest unassigned element.  On a structured box of hexahedra the patches come out as
// compact blocks (about 1.56 patch nodes per element at 512 elements, against 1.42 for a perfect
// 8^3 cube and 8 for no sharing).
static_assert(ElementPatches::kP == kPatchElements, "patch size of the host partition and the kernels");
const ElementPatches *mesh_patches(const sfem_amd_mesh *m) {
    if (m->patches) return m->patches.get();
    if (m->type == SFEM_AMD_SSHEX8 || m->E <= 0 || m->nv <= 0) return nullptr;
    constexpr int P = ElementPatches::kP;
    const int nv = m->nv;
    const ptrdiff_t E = m->E, N = m->nnodes;
    if ((double)E * nv >= 2147483647.0 || N >= 2147483647) {
        set_last_error("element patches: mesh exceeds int32 indexing");
        return nullptr;
    }
    std::vector<int32_t> el((size_t)E * nv);  // SoA [nv][E]
    if (hipMemcpy(el.data(), m->idx.p, sizeof(int32_t) * el.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        set_last_error("element patches: connectivity download failed");
        return nullptr;
    }
    // node -> elements
    std::vector<int32_t> n2e_off((size_t)N + 1, 0);
    for (const int32_t g : el) n2e_off[(size_t)g + 1]++;
    for (ptrdiff_t g = 0; g < N; g++) n2e_off[g + 1] += n2e_off[g];
    std::vector<int32_t> n2e((size_t)E * nv), fill(n2e_off.begin(), n2e_off.end() - 1);
    for (int v = 0; v < nv; v++)
        for (ptrdiff_t e = 0; e < E; e++) n2e[fill[el[(size_t)v * E + e]]++] = (int32_t)e;
    // grow the patches
    std::vector<int32_t> patch_of((size_t)E, -1), order, pstart, queue, hint;
    order.reserve(E);
    ptrdiff_t seed = 0;
    for (;;) {
        int32_t s0 = -1;
        while (!hint.empty()) {
            const int32_t c = hint.back();
            hint.pop_back();
            if (patch_of[c] < 0) {
                s0 = c;
                break;
            }
        }
        if (s0 < 0) {
            while (seed < E && patch_of[seed] >= 0) seed++;
            if (seed >= E) break;
            s0 = (int32_t)seed;
        }
        const int32_t pid = (int32_t)pstart.size();
        pstart.push_back((int32_t)order.size());
        queue.assign(1, s0);
        patch_of[s0] = pid;
        for (size_t head = 0; head < queue.size(); head++) {
            const int32_t e = queue[head];
            for (int v = 0; v < nv && queue.size() < (size_t)P; v++) {
                const int32_t g = el[(size_t)v * E + e];
                for (int32_t k = n2e_off[g]; k < n2e_off[g + 1] && queue.size() < (size_t)P; k++) {
                    const int32_t f = n2e[k];
                    if (patch_of[f] < 0) {
                        patch_of[f] = pid;
                        queue.push_back(f);
                    }
                }
            }
        }
        order.insert(order.end(), queue.begin(), queue.end());
        // unassigned neighbours of this patch, lowest id popped first
        hint.clear();
        for (const int32_t e : queue)
            for (int v = 0; v < nv; v++) {
                const int32_t g = el[(size_t)v * E + e];
                for (int32_t k = n2e_off[g]; k < n2e_off[g + 1]; k++)
                    if (patch_of[n2e[k]] < 0) hint.push_back(n2e[k]);
            }
        std::sort(hint.begin(), hint.end(), std::greater<int32_t>());
        hint.erase(std::unique(hint.begin(), hint.end()), hint.end());
    }
    const ptrdiff_t np = (ptrdiff_t)pstart.size();
    pstart.push_back((int32_t)E);
    // local numbering, patch node lists and which nodes other patches share
    std::vector<int32_t> count((size_t)np), perm((size_t)np * P, -1), node_off(1, 0), nodes;
    std::vector<uint16_t> lconn((size_t)np * nv * P, 0);
    std::vector<int32_t> local((size_t)N, -1), inside((size_t)N, 0);
    std::vector<uint8_t> shared_flag;
    int maxloc = 0;
    for (ptrdiff_t p = 0; p < np; p++) {
        const int32_t b = pstart[p], cnt = pstart[p + 1] - b;
        count[p] = cnt;
        const size_t first = nodes.size();
        for (int32_t t = 0; t < cnt; t++) {
            const int32_t e = order[(size_t)b + t];
            perm[(size_t)p * P + t] = e;
            for (int v = 0; v < nv; v++) {
                const int32_t g = el[(size_t)v * E + e];
                if (local[g] < 0) {
                    local[g] = (int32_t)(nodes.size() - first);
                    nodes.push_back(g);
                }
                inside[g]++;
                lconn[((size_t)p * nv + v) * P + t] = (uint16_t)local[g];
            }
        }
        const int nloc = (int)(nodes.size() - first);
        if (nloc > 65535) {
            set_last_error("element patches: more than 65535 nodes in a patch");
            return nullptr;
        }
        maxloc = std::max(maxloc, nloc);
        for (size_t i = first; i < nodes.size(); i++) {
            const int32_t g = nodes[i];
            shared_flag.push_back(inside[g] < n2e_off[g + 1] - n2e_off[g] ? 1 : 0);
            local[g] = -1;
            inside[g] = 0;
        }
        node_off.push_back((int32_t)nodes.size());
    }
    // node-sorted partial positions of the shared entries (patch order within a node)
    std::vector<int32_t> pcount((size_t)N + 1, 0);
    for (size_t i = 0; i < nodes.size(); i++)
        if (shared_flag[i]) pcount[(size_t)nodes[i] + 1]++;
    std::vector<int32_t> bnode, boff(1, 0), start((size_t)N, -1);
    for (ptrdiff_t g = 0; g < N; g++)
        if (pcount[g + 1]) {
            start[g] = boff.back();
            bnode.push_back((int32_t)g);
            boff.push_back(boff.back() + pcount[g + 1]);
        }
    std::vector<int32_t> bslot(nodes.size(), -1);
    for (size_t i = 0; i < nodes.size(); i++)
        if (shared_flag[i]) bslot[i] = start[nodes[i]]++;
    auto pt = std::make_unique<ElementPatches>();
    pt->npatch = np;
    pt->maxloc = maxloc;
    pt->nodes_per_element = (double)nodes.size() / (double)E;
    pt->nparts = boff.back();
    pt->boundary.n = (ptrdiff_t)bnode.size();
    if (pt->count.upload(count.data(), count.size()) || pt->perm.upload(perm.data(), perm.size()) ||
        pt->lconn.upload(lconn.data(), lconn.size()) || pt->node_off.upload(node_off.data(), node_off.size()) ||
        pt->nodes.upload(nodes.data(), nodes.size()) || pt->bslot.upload(bslot.data(), bslot.size()) ||
        pt->boundary.node.upload(bnode.data(), bnode.size()) || pt->boundary.off.upload(boff.data(), boff.size()))
        return nullptr;
    m->patches = std::move(pt);
    return m->patches.get();
}

// ----------------------------------------------------------------- Laplacian
template <typename T>
struct LaplacianOp final : OpBase {
    const sfem_amd_mesh *m{nullptr};  // the mesh outlives the operator (documented contract)
    DevArray<float> fff32;      // SoA [6][E]
    DevArray<uint16_t> fff16;   // SoA [6][E] bf16 (FP32 path, coefficient folded in)
    DevArray<float> coeff;
    DevArray<T> part;  // sshex8: skeleton contributions [E][sk_count(L)]
    DevArray<double> Kem;  // sshex8 stored sub-element matrices [E][64] (instead of the FFF)
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return m->nnodes; }

    int apply(const void *xv, void *yv, hipStream_t s) override { return apply_impl(xv, yv, s, 0); }
    int apply_overwrite(const void *xv, void *yv, hipStream_t s) override {
        if (m->type != SFEM_AMD_SSHEX8) return OpBase::apply_overwrite(xv, yv, s);
        if (mesh_skeleton(m, 0)) return SFEM_FAILURE;
        if (!skeleton_covers(m)) return OpBase::apply_overwrite(xv, yv, s);
        return apply_impl(xv, yv, s, 1);
    }
    const sfem_amd_mesh *gather_mesh(hipStream_t) override { return m->type == SFEM_AMD_SSHEX8 ? m : nullptr; }
    int apply_constrained(const void *xv, void *yv, const uint8_t *cmask, int ow, hipStream_t s) override {
        if (m->type != SFEM_AMD_SSHEX8 || mesh_skeleton(m, 0)) return SFEM_FAILURE;
        if (ow && !skeleton_covers(m)) {
            if (fill_zero(yv, (size_t)rows() * sizeof(T), s)) return SFEM_FAILURE;
            ow = 0;
        }
        return apply_impl(xv, yv, s, ow, cmask);
    }
    int apply_impl(const void *xv, void *yv, hipStream_t s, int ow, const uint8_t *cmask = nullptr) {
        const T *x = (const T *)xv;
        T *y = (T *)yv;
        if (m->type == SFEM_AMD_SSHEX8 && !part.p && m->E > 0) {
            if (mesh_skeleton(m, 0) || part.alloc((size_t)m->E * sk_count(m->level))) return SFEM_FAILURE;
        }
        SkeletonMap sk = skeleton_of(m);
        sk.cmask = cmask;
        if constexpr (std::is_same<T, double>::value) {
            switch (m->type) {
                case SFEM_AMD_TET4:
                    return tet4_laplacian_apply(m->E, m->idx.p, fff32.p, x, y, s);
                case SFEM_AMD_HEX8:
                    return hex8_laplacian_apply(m->E, m->idx.p, fff32.p, coeff.p, x, y, s);
                default:
                    return sshex8_laplacian_apply<double, float>(m->level, m->E, m->idx.p, fff32.p, m->E, coeff.p,
                                                                 Kem.p, x, y, part.p, sk, s, ow, g_atomic_skeleton);
            }
        } else {
            return sshex8_laplacian_apply<float, uint16_t>(m->level, m->E, m->idx.p, fff16.p, m->E, nullptr, Kem.p,
                                                           x, y, part.p, sk, s, ow, g_atomic_skeleton);
        }
    }
    int diag(void *dv, hipStream_t s) override {
        T *d = (T *)dv;
        if constexpr (std::is_same<T, double>::value) {
            switch (m->type) {
                case SFEM_AMD_TET4:
                    return tet4_laplacian_diag(m->E, m->idx.p, fff32.p, d, s);
                case SFEM_AMD_HEX8:
                    return hex8_laplacian_diag(m->E, m->idx.p, fff32.p, coeff.p, d, s);
                default:
                    return sshex8_laplacian_diag<double, float>(m->level, m->E, m->idx.p, fff32.p, m->E, coeff.p,
                                                                Kem.p, d, s);
            }
        } else {
            return sshex8_laplacian_diag<float, uint16_t>(m->level, m->E, m->idx.p, fff16.p, m->E, nullptr, Kem.p, d,
                                                          s);
        }
    }
};

// ----------------------------------------------------------------- elasticity
template <typename T>
struct ElasticityOp final : OpBase {
    const sfem_amd_mesh *m{nullptr};
    DevArray<float> adj, det, coeff;  // hex8: SoA [9][E], [E]
    DevArray<uint16_t> adj16, det16;  // hex8, FP32 operator: the same words rounded to bf16
    // hex8 on element patches: geometry and coefficient words in patch-slot order, shared-node partials
    DevArray<float> adjp, detp, coeffp;
    DevArray<uint16_t> adjp16, detp16;
    DevArray<T> ppart;
    int patch_state{0};  // 0 not yet tried, 1 ready, -1 not applicable
    DevArray<double> K;               // sshex8: per-macro 24x24 matrices
    DevArray<float> K32;              // sshex8, FP32 operator: K rounded to float
    DevArray<uint32_t> Apk;           // FP32 matrix cores: K32 split into bf16 A operands
    // FP32 from geometry: the bf16x3 MFMA kernel (FP64 always runs on the matrix cores: exact)
    bool matrix_cores{false};
    DevArray<T> part;                 // sshex8: skeleton contributions [E][sk_count(L)][3]
    const sfem_amd_mesh *part_mesh{nullptr};  // the mesh `part` was sized for
    // levels beyond the matrix-core kernels' LDS tile (L = 2 Ls): the operator applied as 8 sub-blocks
    // of level Ls per macro element (sharing its matrix), with their own skeleton slots
    double mu{0}, lambda{0};
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return m->nnodes * 3; }
    bool uses_matrix_cores() const { return std::is_same<T, double>::value || matrix_cores; }
    // the stencil form (27-point 3x3-block stencil per macro element, sshex8_elasticity_apply_stencil):
    // option elasticity_form 2; by default (0) the FP64 operators at levels >= 10, where it beats the
    // element form (L = 16 on a 24^3 macro cube: 2.5 vs 3.0 ms; the FP32 matrix-core element form
    // and the FP64 one at L <= 8 stay faster); 1 keeps the element form
    bool uses_stencil() const {
        if (m->type != SFEM_AMD_SSHEX8 || !sshex8_elasticity_stencil_supported(m->level)) return false;
        if (g_elasticity_form == 1) return false;
        if (g_elasticity_form == 2) return true;
        return std::is_same<T, double>::value && m->level >= 10;
    }
    // the sub-block level, or 0 when the macro elements are applied directly
    int subblock_level() const {
        if (m->type != SFEM_AMD_SSHEX8 || !uses_matrix_cores() || uses_stencil()) return 0;
        const int L = m->level;
        if constexpr (std::is_same<T, double>::value) {
            if (L % 2 == 0 && L <= 8) return 0;
            return (L == 12 || L == 16) ? L / 2 : 0;
        } else {
            if (L % 2 == 0 && L <= 12) return 0;
            return L == 16 ? 8 : 0;
        }
    }
    // the mesh the kernels run on (built on first use)
    const sfem_amd_mesh *work_mesh(hipStream_t s) {
        const int Ls = subblock_level();
        return Ls ? mesh_subblocks(m, Ls, s) : m;
    }
    int apply(const void *xv, void *yv, hipStream_t s) override { return apply_impl(xv, yv, s, 0); }
    int apply_overwrite(const void *xv, void *yv, hipStream_t s) override {
        if (m->type != SFEM_AMD_SSHEX8) return OpBase::apply_overwrite(xv, yv, s);
        const sfem_amd_mesh *wm = work_mesh(s);
        if (!wm || mesh_skeleton(wm, 0)) return SFEM_FAILURE;
        if (!skeleton_covers(wm)) return OpBase::apply_overwrite(xv, yv, s);
        return apply_impl(xv, yv, s, 1);
    }
    int block_size() const override { return 3; }
    const sfem_amd_mesh *gather_mesh(hipStream_t s) override {
        return m->type == SFEM_AMD_SSHEX8 ? work_mesh(s) : nullptr;
    }
    int apply_constrained(const void *xv, void *yv, const uint8_t *cmask, int ow, hipStream_t s) override {
        const sfem_amd_mesh *wm = m->type == SFEM_AMD_SSHEX8 ? work_mesh(s) : nullptr;
        if (!wm || mesh_skeleton(wm, 0)) return SFEM_FAILURE;
        if (ow && !skeleton_covers(wm)) {
            if (fill_zero(yv, (size_t)rows() * sizeof(T), s)) return SFEM_FAILURE;
            ow = 0;
        }
        return apply_impl(xv, yv, s, ow, cmask);
    }
    bool cheb3_fusable(hipStream_t s) override {
        if (!g_fused_chebyshev || m->type != SFEM_AMD_SSHEX8 || uses_stencil()) return false;
        const sfem_amd_mesh *wm = work_mesh(s);
        if (!wm || mesh_skeleton(wm, 0) || !skeleton_covers(wm)) return false;
        return sshex8_elasticity_cheb3_fusable(wm->level, std::is_same<T, double>::value, uses_matrix_cores());
    }
    int cheb3_step(void *x, const void *rhs, const void *dinv, void *p, double alpha, double beta, int first,
                   const uint8_t *cmask, hipStream_t s) override {
        ChebStep<T> cs;
        cs.rhs = (const T *)rhs;
        cs.dinv = (const T *)dinv;
        cs.p = (T *)p;
        cs.x = (T *)x;
        cs.alpha = (T)alpha;
        cs.beta = (T)beta;
        cs.first = first;
        return apply_impl(x, p, s, 1, cmask, &cs);  // y is not touched by the fused step
    }
    int apply_impl(const void *xv, void *yv, hipStream_t s, int ow, const uint8_t *cmask = nullptr,
                   const ChebStep<T> *cs = nullptr) {
        if (m->type == SFEM_AMD_HEX8) {
            if (g_hex8_patches && m->E >= kPatchMinElements && patches_ready(s)) return apply_patches(xv, yv, s);
            if constexpr (std::is_same<T, double>::value)
                return hex8_elasticity_apply<double, float>(m->E, m->idx.p, adj.p, det.p, mu, lambda, coeff.p,
                                                            (const double *)xv, (double *)yv, s);
            else
                return hex8_elasticity_apply<float, uint16_t>(m->E, m->idx.p, adj16.p, det16.p, mu, lambda, coeff.p,
                                                              (const float *)xv, (float *)yv, s);
        }
        const sfem_amd_mesh *wm = work_mesh(s);
        if (!wm) return SFEM_FAILURE;
        // the partials are sized for the mesh the kernels run on (the form may change with the option)
        if ((!part.p || part_mesh != wm) && wm->E > 0) {
            // (+3: the stencil kernel's sink for lanes without a store)
            if (mesh_skeleton(wm, 0) || part.alloc((size_t)wm->E * sk_count(wm->level) * 3 + 3)) return SFEM_FAILURE;
            part_mesh = wm;
        }
        const bool stencil = uses_stencil() && !cs;
        const T *Kt;
        if constexpr (std::is_same<T, double>::value) {
            Kt = K.p;
        } else {
            if (!K32.p && K.n) {
                if (K32.alloc(K.n) || blas_cast<double, float>((ptrdiff_t)K.n, K.p, K32.p, s)) return SFEM_FAILURE;
            }
            Kt = K32.p;
            if (matrix_cores && !stencil && !Apk.p && K32.n) {
                const ptrdiff_t nK = (ptrdiff_t)K32.n / 576;
                if (Apk.alloc((size_t)nK * 1024) || sshex8_elasticity_pack_bf16x2(nK, K32.p, Apk.p, s))
                    return SFEM_FAILURE;
            }
        }
        SkeletonMap sk = skeleton_of(wm);
        sk.cmask = cmask;
        if (stencil)
            return sshex8_elasticity_apply_stencil<T>(m->level, m->E, m->idx.p, K.p, (const T *)xv, (T *)yv, part.p,
                                                      sk, s, ow);
        return sshex8_elasticity_apply<T>(wm->level, wm->E, wm->idx.p, Kt, (const T *)xv, (T *)yv, part.p, sk, s, ow,
                                          uses_matrix_cores() ? 1 : 0, wm == m ? 0 : 3, Apk.p,
                                          cs ? *cs : ChebStep<T>());
    }
    // hex8: the element patches of the mesh and this operator's words in their slot order (first use)
    bool patches_ready(hipStream_t s) {
        if (patch_state) return patch_state > 0;
        patch_state = -1;
        const ElementPatches *pt = mesh_patches(m);
        if (!pt || hex8_patch_lds_bytes(pt->maxloc, sizeof(T)) > kPatchLdsLimit) return false;
        const ptrdiff_t Ep = pt->npatch * ElementPatches::kP;
        bool ok = ppart.alloc((size_t)pt->nparts * 3 + 3) == SFEM_SUCCESS;
        if constexpr (std::is_same<T, double>::value)
            ok = ok && !adjp.alloc((size_t)9 * Ep) && !detp.alloc(Ep) &&
                 !patch_gather_rows<float>(9, m->E, Ep, pt->perm.p, adj.p, adjp.p, s) &&
                 !patch_gather_rows<float>(1, m->E, Ep, pt->perm.p, det.p, detp.p, s);
        else
            ok = ok && !adjp16.alloc((size_t)9 * Ep) && !detp16.alloc(Ep) &&
                 !patch_gather_rows<uint16_t>(9, m->E, Ep, pt->perm.p, adj16.p, adjp16.p, s) &&
                 !patch_gather_rows<uint16_t>(1, m->E, Ep, pt->perm.p, det16.p, detp16.p, s);
        if (ok && coeff.p)
            ok = !coeffp.alloc(Ep) && !patch_gather_rows<float>(1, m->E, Ep, pt->perm.p, coeff.p, coeffp.p, s);
        if (!ok || hipStreamSynchronize(s) != hipSuccess) return false;
        patch_state = 1;
        return true;
    }
    int apply_patches(const void *xv, void *yv, hipStream_t s) {
        const ElementPatches *pt = m->patches.get();
        PatchView pv;
        pv.npatch = pt->npatch;
        pv.maxloc = pt->maxloc;
        pv.count = pt->count.p;
        pv.node_off = pt->node_off.p;
        pv.nodes = pt->nodes.p;
        pv.bslot = pt->bslot.p;
        pv.lconn = pt->lconn.p;
        int err;
        if constexpr (std::is_same<T, double>::value)
            err = hex8_elasticity_apply_patches<double, float>(pv, adjp.p, detp.p, coeffp.p, mu, lambda,
                                                               (const double *)xv, (double *)yv, ppart.p, s);
        else
            err = hex8_elasticity_apply_patches<float, uint16_t>(pv, adjp16.p, detp16.p, coeffp.p, mu, lambda,
                                                                 (const float *)xv, (float *)yv, ppart.p, s);
        if (err) return err;
        SkeletonMap sk{};
        sk.node = pt->boundary.node.p;
        sk.off = pt->boundary.off.p;
        sk.n = pt->boundary.n;
        return skeleton_gather<T>(sk, 3, ppart.p, (T *)yv, s, 0, 1);
    }
    int hex8_diag(int block, void *d, hipStream_t s) {
        if constexpr (std::is_same<T, double>::value)
            return hex8_elasticity_diag<double, float>(m->E, m->idx.p, adj.p, det.p, mu, lambda, coeff.p, block,
                                                       (double *)d, s);
        else
            return hex8_elasticity_diag<float, uint16_t>(m->E, m->idx.p, adj16.p, det16.p, mu, lambda, coeff.p,
                                                         block, (float *)d, s);
    }
    int diag(void *d, hipStream_t s) override {
        if (m->type == SFEM_AMD_HEX8) return hex8_diag(0, d, s);
        return sshex8_elasticity_diag<T>(m->level, m->E, m->idx.p, K.p, 0, (T *)d, s);
    }
    int block_diag(void *d, hipStream_t s) override {
        if (m->type == SFEM_AMD_HEX8) return hex8_diag(1, d, s);
        return sshex8_elasticity_diag<T>(m->level, m->E, m->idx.p, K.p, 1, (T *)d, s);
    }
};

// ----------------------------------------------------------------- constraints
// The dofs (node * bs + component) as bits on the skeleton entries of sk; fails when a dof's node
// is not a skeleton node (or bs > 8).
static int skeleton_dof_mask(const SkeletonMap &sk, int bs, const DevArray<int32_t> &dofs, DevArray<uint8_t> &mask_out,
                             hipStream_t s) {
    if (bs > 8 || sk.n == 0) return SFEM_FAILURE;
    std::vector<int32_t> node((size_t)sk.n), d(dofs.n);
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(node.data(), sk.node, sizeof(int32_t) * node.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        (!d.empty() && hipMemcpy(d.data(), dofs.p, sizeof(int32_t) * d.size(), hipMemcpyDeviceToHost) != hipSuccess))
        return SFEM_FAILURE;
    std::vector<uint8_t> mask((size_t)sk.n, 0);
    for (const int32_t q : d) {
        const int32_t g = q / bs;
        const auto it = std::lower_bound(node.begin(), node.end(), g);  // skeleton nodes ascend
        if (it == node.end() || *it != g) return SFEM_FAILURE;         // an interior dof
        mask[(size_t)(it - node.begin())] |= (uint8_t)(1u << (q % bs));
    }
    return mask_out.upload(mask.data(), mask.size());
}

struct ConstrainedOp final : OpBase {
    std::shared_ptr<OpBase> inner;
    DevArray<int32_t> dofs;
    // the constrained dofs as flags on the inner sshex8 operator's skeleton entries, when every
    // one is a skeleton (macro-element boundary) dof: its gather then writes y = x there
    DevArray<uint8_t> cmask;
    int cmask_state{0};  // 0: not yet built, 1: in use, -1: not applicable (separate copy)
    bool fused(hipStream_t s) {
        if (cmask_state == 0) {
            cmask_state = -1;
            if (!build_cmask(s)) cmask_state = 1;
            (void)hipGetLastError();
        }
        return cmask_state > 0;
    }
    int build_cmask(hipStream_t s) {
        const sfem_amd_mesh *gm = inner->gather_mesh(s);
        if (!gm || mesh_skeleton(gm, 0)) return SFEM_FAILURE;
        return skeleton_dof_mask(skeleton_of(gm), inner->block_size(), dofs, cmask, s);
    }
    int apply(const void *x, void *y, hipStream_t s) override {
        if (fused(s)) return inner->apply_constrained(x, y, cmask.p, 0, s);
        if (inner->apply(x, y, s)) return SFEM_FAILURE;
        return copy_constrained(x, y, s);
    }
    int apply_overwrite(const void *x, void *y, hipStream_t s) override {
        if (fused(s)) return inner->apply_constrained(x, y, cmask.p, 1, s);
        if (inner->apply_overwrite(x, y, s)) return SFEM_FAILURE;
        return copy_constrained(x, y, s);
    }
    bool cheb3_fusable(hipStream_t s) override { return fused(s) && inner->cheb3_fusable(s); }
    int cheb3_step(void *x, const void *rhs, const void *dinv, void *p, double alpha, double beta, int first,
                   const uint8_t *, hipStream_t s) override {
        return inner->cheb3_step(x, rhs, dinv, p, alpha, beta, first, cmask.p, s);
    }
    int copy_constrained(const void *x, void *y, hipStream_t s) const {
        if (inner->real_type() == SFEM_AMD_FLOAT64)
            return dofs_copy<double>((ptrdiff_t)dofs.n, dofs.p, (const double *)x, (double *)y, s);
        return dofs_copy<float>((ptrdiff_t)dofs.n, dofs.p, (const float *)x, (float *)y, s);
    }
    // x[c] = v for a vector of real type rt (which may differ from the operator's)
    int set_value_rt(int rt, double v, void *x, hipStream_t s) const {
        if (rt == SFEM_AMD_FLOAT64) return dofs_set_value<double>((ptrdiff_t)dofs.n, dofs.p, v, (double *)x, s);
        return dofs_set_value<float>((ptrdiff_t)dofs.n, dofs.p, (float)v, (float *)x, s);
    }
    int set_value(double v, void *x, hipStream_t s) const { return set_value_rt(inner->real_type(), v, x, s); }
    ptrdiff_t rows() const override { return inner->rows(); }
    int real_type() const override { return inner->real_type(); }
    int diag(void *d, hipStream_t s) override {
        if (inner->diag(d, s)) return SFEM_FAILURE;
        return set_value(1.0, d, s);  // f->set_value_to_constrained_dofs(1, diag)
    }
    int block_diag(void *d, hipStream_t s) override { return inner->block_diag(d, s); }
};

// ----------------------------------------------------------------- transfers
template <typename T>
struct TransferOp final : OpBase {
    const sfem_amd_mesh *fine{nullptr};
    int Lc{1};
    ptrdiff_t n_coarse{0};
    int vs{1};
    bool prolong{true};
    std::shared_ptr<OpBase> constrained;  // ConstrainedOp on the output level (or null)
    DevArray<T> part;                     // restriction: coarse skeleton contributions
    DevArray<uint8_t> keep;               // adding prolongation: constrained fine dofs (bit per component)
    bool keep_built{false};
    DevArray<uint8_t> rmask;              // restriction: constrained coarse dofs on the coarse skeleton
    int rmask_state{0};                   // 0 not built, 1 the gather zeroes them, -1 separate zeroing
    // the element tile fits the default LDS budget: use the LDS-staged kernels
    bool staged() const {
        return (vs == 1 || vs == 3) && sshex8_transfer_lds_bytes(prolong ? Lc : fine->level, vs) <= 48 * 1024;
    }
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return prolong ? fine->nnodes * vs : n_coarse * vs; }
    ptrdiff_t cols() const override { return prolong ? n_coarse * vs : fine->nnodes * vs; }
    bool adds_correction(hipStream_t) override { return prolong && staged(); }
    // y += P x, constrained fine dofs unchanged: the same words as zero + apply (overwriting) +
    // zeroing the constrained rows + y += that
    int apply_add(const void *x, void *y, hipStream_t s) override {
        if (!prolong || !staged()) return OpBase::apply_add(x, y, s);
        if (mesh_owner(fine, s)) return SFEM_FAILURE;
        if (!keep_built) {
            if (constrained) {
                auto *c = static_cast<ConstrainedOp *>(constrained.get());
                std::vector<int32_t> d(c->dofs.n);
                if (!d.empty() && hipMemcpy(d.data(), c->dofs.p, sizeof(int32_t) * d.size(), hipMemcpyDeviceToHost) !=
                                          hipSuccess) {
                    set_last_error("prolongation: reading the constrained dofs failed");
                    return SFEM_FAILURE;
                }
                std::vector<uint8_t> k((size_t)fine->nnodes, 0);
                for (const int32_t q : d)
                    if (q >= 0 && q / vs < fine->nnodes) k[(size_t)(q / vs)] |= (uint8_t)(1u << (q % vs));
                if (keep.upload(k.data(), k.size())) return SFEM_FAILURE;
            }
            keep_built = true;
        }
        return sshex8_prolongate_lds<T>(fine->E, fine->level, Lc, fine->idx.p, fine->own.p, vs, (const T *)x, (T *)y,
                                        s, keep.p, 1);
    }
    int apply_overwrite(const void *x, void *y, hipStream_t s) override {
        if (prolong) {
            if (!staged()) return OpBase::apply_overwrite(x, y, s);
            return apply(x, y, s);  // every fine node is written by its owner element
        }
        return restrict_impl(x, y, s, 1);
    }
    int apply(const void *x, void *y, hipStream_t s) override {
        if (!prolong) return restrict_impl(x, y, s, 0);
        int err;
        {
            if (mesh_owner(fine, s)) return SFEM_FAILURE;
            if (staged())
                err = sshex8_prolongate_lds<T>(fine->E, fine->level, Lc, fine->idx.p, fine->own.p, vs, (const T *)x,
                                               (T *)y, s);
            else
                err = sshex8_prolongate<T>(fine->E, fine->level, Lc, fine->idx.p, fine->own.p, vs, (const T *)x,
                                           (T *)y, s);
        }
        if (err) return err;
        if (constrained) return static_cast<ConstrainedOp *>(constrained.get())->set_value_rt(real_type(), 0.0, y, s);
        return SFEM_SUCCESS;
    }
    int restrict_impl(const void *x, void *y, hipStream_t s, int ow) {
        int err;
        {
            // a fine element tile beyond the LDS budget: restrict its 2^3 sub-blocks of level L/2 onto
            // level Lc/2 instead (the same operator: the sub-blocks align with coarse micro elements)
            const sfem_amd_mesh *rf = fine;
            int rlc = Lc;
            if (!staged() && (vs == 1 || vs == 3) && fine->level % 2 == 0 && Lc % 2 == 0 &&
                sshex8_transfer_lds_bytes(fine->level / 2, vs) <= 48 * 1024) {
                rf = mesh_subblocks(fine, fine->level / 2, s);
                if (!rf) return SFEM_FAILURE;
                rlc = Lc / 2;
            }
            if (mesh_counts(rf, s)) return SFEM_FAILURE;
            if (rf != fine || staged()) {
                if (mesh_skeleton(rf, rlc)) return SFEM_FAILURE;
                if (!part.p && part.alloc((size_t)rf->E * sk_count(rlc) * vs)) return SFEM_FAILURE;
                SkeletonMap sk = skeleton_of(rf, rlc);
                if (constrained && rmask_state == 0) {
                    rmask_state = skeleton_dof_mask(sk, vs, static_cast<ConstrainedOp *>(constrained.get())->dofs,
                                                    rmask, s)
                                          ? -1
                                          : 1;
                    (void)hipGetLastError();
                }
                if (rmask_state > 0) sk.cmask = rmask.p;  // the gather writes 0 at constrained coarse dofs
                err = sshex8_restrict_lds<T>(rf->E, rf->level, rlc, rf->idx.p, rf->count.p, vs, (const T *)x, (T *)y,
                                             part.p, sk.pos, s, ow);
                if (!err) err = skeleton_gather<T>(sk, vs, part.p, (T *)y, s, ow);
                if (!err && rmask_state > 0) return SFEM_SUCCESS;
            } else {
                if (ow && fill_zero(y, (size_t)rows() * sizeof(T), s)) return SFEM_FAILURE;
                err = sshex8_restrict<T>(fine->E, fine->level, Lc, fine->idx.p, fine->count.p, vs, (const T *)x,
                                         (T *)y, s);
            }
        }
        if (err) return err;
        if (constrained) return static_cast<ConstrainedOp *>(constrained.get())->set_value_rt(real_type(), 0.0, y, s);
        return SFEM_SUCCESS;
    }
};

// ----------------------------------------------------------------- permutations and chains
template <typename T>
struct PermutationOp final : OpBase {
    DevArray<int32_t> perm;
    ptrdiff_t n{0};
    int bs{1}, transpose{0};
    std::shared_ptr<OpBase> constrained;
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return n * bs; }
    int apply(const void *x, void *y, hipStream_t s) override {
        if (permute_add<T>(n, bs, perm.p, transpose, (const T *)x, (T *)y, s)) return SFEM_FAILURE;
        if (constrained) return static_cast<ConstrainedOp *>(constrained.get())->set_value_rt(real_type(), 0.0, y, s);
        return SFEM_SUCCESS;
    }
};

struct ChainOp final : OpBase {
    std::shared_ptr<OpBase> A, B;
    DevArray<char> tmp;
    int real_type() const override { return A->real_type(); }
    ptrdiff_t rows() const override { return B->rows(); }
    ptrdiff_t cols() const override { return A->cols(); }
    int apply(const void *x, void *y, hipStream_t s) override {
        const size_t bytes = (size_t)A->rows() * (real_type() == SFEM_AMD_FLOAT64 ? 8 : 4);
        if (tmp.n != bytes && tmp.alloc(bytes)) return SFEM_FAILURE;
        if (fill_zero(tmp.p, bytes, s)) return SFEM_FAILURE;
        if (A->apply(x, tmp.p, s)) return SFEM_FAILURE;
        return B->apply(tmp.p, y, s);
    }
};

// ----------------------------------------------------------------- sfem:: template bridges
// A runtime operator seen through the sfem:: interfaces.  It is a MatrixFreeLinearSolver so that
// a multigrid or CG can pass set_initial_guess_zero on to a wrapped smoother or preconditioner;
// the solver set-up calls are not forwarded (the wrapped object is configured at creation).
template <typename T>
struct OpView final : sfem::MatrixFreeLinearSolver<T>, sfem::FusedChebyshevStep<T>, sfem::AddingProlongation<T> {
    std::shared_ptr<OpBase> impl;
    hipStream_t s{nullptr};
    bool overwrite{false};  // apply() overwrites y (the owning solver skips its zero fill)
    explicit OpView(std::shared_ptr<OpBase> o, bool ow = false) : impl(std::move(o)), overwrite(ow) {}
    int apply(const T *const x, T *const y) override {
        return overwrite ? impl->apply_overwrite(x, y, as_stream(s)) : impl->apply(x, y, as_stream(s));
    }
    void set_op(const sfem::SharedOperator<T> &) override {}
    void set_preconditioner_op(const sfem::SharedOperator<T> &) override {}
    void set_max_it(int) override {}
    void set_n_dofs(std::ptrdiff_t) override {}
    void set_initial_guess_zero(bool v) override { impl->set_initial_guess_zero(v); }
    bool zero_guess_ignores_x() const override { return impl->zero_guess_ignores_x(); }
    int iterations() const override { return impl->iterations(); }
    bool cheb3_fusable() override { return impl->cheb3_fusable(as_stream(s)); }
    bool adds_correction() override { return impl->adds_correction(as_stream(s)); }
    int apply_add(const T *coarse, T *fine) override { return impl->apply_add(coarse, fine, as_stream(s)); }
    int cheb3_step(T *x, const T *rhs, const T *inv_diag, T *p, T alpha, T beta, int first) override {
        return impl->cheb3_step(x, rhs, inv_diag, p, alpha, beta, first, nullptr, as_stream(s));
    }
    std::ptrdiff_t rows() const override { return impl->rows(); }
    std::ptrdiff_t cols() const override { return impl->cols(); }
    sfem::ExecutionSpace execution_space() const override { return sfem::EXECUTION_SPACE_DEVICE; }
};

template <typename T>
static sfem::SharedOperator<T> view(const std::shared_ptr<OpBase> &o, bool overwrite = false) {
    if (!o) return nullptr;
    return std::make_shared<OpView<T>>(o, overwrite);
}

template <typename T>
struct SolverOp final : OpBase {
    std::shared_ptr<sfem::Operator<T>> op;
    std::function<int()> its;
    std::function<int()> status;
    std::function<const std::vector<double> *()> hist;
    std::vector<std::shared_ptr<OpBase>> keep;  // children kept alive
    ptrdiff_t n{0};
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return n; }
    int apply(const void *x, void *y, hipStream_t) override { return op->apply((const T *)x, (T *)y); }
    void set_initial_guess_zero(bool v) override {
        if (auto *m = dynamic_cast<sfem::MatrixFreeLinearSolver<T> *>(op.get()))
            m->set_initial_guess_zero(v);
        else if (auto *mg = dynamic_cast<sfem::Multigrid<T> *>(op.get()))
            mg->set_initial_guess_zero(v);
    }
    bool zero_guess_ignores_x() const override {
        auto *m = dynamic_cast<const sfem::MatrixFreeLinearSolver<T> *>(op.get());
        return m && m->zero_guess_ignores_x();
    }
    int iterations() const override { return its ? its() : -1; }
    int last_status() const override { return status ? status() : SFEM_SUCCESS; }
    const std::vector<double> *history() const override { return hist ? hist() : nullptr; }
};

// FP64 -> FP32 operator -> FP64: y += (double) op32((float) x)
struct PrecisionBridge final : OpBase {
    std::shared_ptr<OpBase> op32;
    DevArray<float> xf, yf;
    int real_type() const override { return SFEM_AMD_FLOAT64; }
    ptrdiff_t rows() const override { return op32->rows(); }
    int apply(const void *x, void *y, hipStream_t s) override {
        const ptrdiff_t n = rows();
        if ((ptrdiff_t)xf.n != n && (xf.alloc(n) || yf.alloc(n))) return SFEM_FAILURE;
        if (blas_cast<double, float>(n, (const double *)x, xf.p, s)) return SFEM_FAILURE;
        if (blas_values<float>(n, 0.f, yf.p, s)) return SFEM_FAILURE;
        op32->set_initial_guess_zero(true);  // yf was just zeroed
        const int err = op32->apply(xf.p, yf.p, s);
        op32->set_initial_guess_zero(false);
        if (err) return SFEM_FAILURE;
        return cast_accumulate(n, yf.p, (double *)y, s);  // y += (double) yf
    }
    // y = (double) op32((float) x): the result is cast into y, which is neither cleared nor read
    int apply_overwrite(const void *x, void *y, hipStream_t s) override {
        const ptrdiff_t n = rows();
        if ((ptrdiff_t)xf.n != n && (xf.alloc(n) || yf.alloc(n))) return SFEM_FAILURE;
        if (blas_cast<double, float>(n, (const double *)x, xf.p, s)) return SFEM_FAILURE;
        if (blas_values<float>(n, 0.f, yf.p, s)) return SFEM_FAILURE;
        op32->set_initial_guess_zero(true);
        const int err = op32->apply(xf.p, yf.p, s);
        op32->set_initial_guess_zero(false);
        if (err) return SFEM_FAILURE;
        return blas_cast<float, double>(n, yf.p, (double *)y, s);
    }
};

// FP32 -> FP64 operator -> FP32: y += (float) op64((double) x).  Used for the FP64
// coarse solve inside an FP32 multigrid (x: coarse rhs, y: zeroed coarse solution).
struct PrecisionBridgeDown final : OpBase {
    std::shared_ptr<OpBase> op64;
    DevArray<double> xd, yd;
    int real_type() const override { return SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return op64->rows(); }
    ptrdiff_t cols() const override { return op64->cols(); }
    int apply(const void *x, void *y, hipStream_t s) override {
        const ptrdiff_t n = rows(), m = cols();
        if ((ptrdiff_t)xd.n != m && xd.alloc(m)) return SFEM_FAILURE;
        if ((ptrdiff_t)yd.n != n && yd.alloc(n)) return SFEM_FAILURE;
        if (blas_cast<float, double>(m, (const float *)x, xd.p, s)) return SFEM_FAILURE;
        if (blas_values<double>(n, 0.0, yd.p, s)) return SFEM_FAILURE;
        if (op64->apply(xd.p, yd.p, s)) return SFEM_FAILURE;
        return cast_accumulate_f32(n, yd.p, (float *)y, s);
    }
    int iterations() const override { return op64->iterations(); }
};

// Solver in the other precision with x as initial guess and output.
struct SolverBridge final : OpBase {
    std::shared_ptr<OpBase> inner;
    DevArray<char> bd, xd;
    int real_type() const override { return inner->real_type() == SFEM_AMD_FLOAT64 ? SFEM_AMD_FLOAT32 : SFEM_AMD_FLOAT64; }
    ptrdiff_t rows() const override { return inner->rows(); }
    int apply(const void *b, void *x, hipStream_t s) override {
        const ptrdiff_t n = rows();
        const int rin = inner->real_type(), rout = real_type();
        const size_t bytes = (size_t)n * (rin == SFEM_AMD_FLOAT64 ? 8 : 4);
        if (bd.n != bytes && (bd.alloc(bytes) || xd.alloc(bytes))) return SFEM_FAILURE;
        if (sfem_amd_blas_cast(rout, n, b, rin, bd.p, s) || sfem_amd_blas_cast(rout, n, x, rin, xd.p, s))
            return SFEM_FAILURE;
        if (inner->apply(bd.p, xd.p, s)) return SFEM_FAILURE;
        return sfem_amd_blas_cast(rin, n, xd.p, rout, x, s);
    }
    int iterations() const override { return inner->iterations(); }
};

}  // namespace sfem_amd

using namespace sfem_amd;

static sfem_amd_op *wrap(std::shared_ptr<OpBase> o) {
    auto *h = new sfem_amd_op;
    h->impl = std::move(o);
    return h;
}

namespace sfem_amd {
sfem_amd_op *wrap_op(std::shared_ptr<OpBase> o) { return wrap(std::move(o)); }
int constrained_set_value(const OpBase *constrained, int rt, double v, void *x, hipStream_t s) {
    auto *c = dynamic_cast<const ConstrainedOp *>(constrained);
    if (!c) {
        set_last_error("not a constrained operator");
        return SFEM_FAILURE;
    }
    return c->set_value_rt(rt, v, x, s);
}
bool is_constrained(const OpBase *op) { return dynamic_cast<const ConstrainedOp *>(op) != nullptr; }
}  // namespace sfem_amd

namespace {
// y += M x with a dense FP64 matrix (coarsest-level direct solve through a precomputed inverse).
template <typename T>
struct DenseOp final : OpBase {
    DevArray<double> M;
    ptrdiff_t n{0};
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return n; }
    int apply(const void *x, void *y, hipStream_t s) override {
        return dense_matvec<T>(n, M.p, (const T *)x, (T *)y, s);
    }
};

// Replays the apply of `inner` as a HIP graph.  Per initial-guess mode (set_initial_guess_zero,
// which changes the recorded sequence) the first call runs eagerly (lazy allocations, one-time
// kernel attributes), the second captures the launch sequence on its stream and launches the
// graph; later calls with the same (x, y, stream) and mode relaunch it, other arguments run
// eagerly.  The inner apply must not synchronise with the host (e.g. a preconditioning V-cycle
// with a direct coarse solve).  A capture that cannot begin leaves the operator eager; an inner
// apply that synchronises inside the capture invalidates it, and the HIP runtime then reports
// that error on the next launch on the stream, so such an apply fails instead of running eagerly.
struct GraphOp final : OpBase {
    std::shared_ptr<OpBase> inner;
    struct Capture {
        hipGraphExec_t exec{nullptr};
        const void *cx{nullptr};
        void *cy{nullptr};
        hipStream_t cs{nullptr};
        int calls{0};
        uint64_t epoch{0};  // the state epoch of the capture
    } cap[4];  // [2 * overwrite + zero initial guess]
    int replays{0};
    bool failed{false};
    bool zero_guess{false};
    ~GraphOp() override {
        for (auto &c : cap)
            if (c.exec) (void)hipGraphExecDestroy(c.exec);
    }
    int captured() const {
        int n = 0;
        for (const auto &c : cap) n += c.exec ? 1 : 0;
        return n;
    }
    int real_type() const override { return inner->real_type(); }
    ptrdiff_t rows() const override { return inner->rows(); }
    ptrdiff_t cols() const override { return inner->cols(); }
    int iterations() const override { return inner->iterations(); }
    void set_initial_guess_zero(bool v) override {
        zero_guess = v;
        inner->set_initial_guess_zero(v);
    }
    int apply(const void *x, void *y, hipStream_t s) override { return run(x, y, s, false); }
    int apply_overwrite(const void *x, void *y, hipStream_t s) override { return run(x, y, s, true); }
    // the inner operator's apply (accumulating) or apply_overwrite, replayed from its own capture
    int run(const void *x, void *y, hipStream_t s, bool overwrite) {
        auto eager = [&] { return overwrite ? inner->apply_overwrite(x, y, s) : inner->apply(x, y, s); };
        Capture &c = cap[(overwrite ? 2 : 0) + (zero_guess ? 1 : 0)];
        const uint64_t now = state_epoch();
        if (c.exec && c.epoch != now) {
            // options, switches or weights changed since the capture: drop it, run the next call
            // eagerly (lazy set-up such as a weight repack happens outside a capture), then capture
            (void)hipGraphExecDestroy(c.exec);
            c.exec = nullptr;
            c.calls = 0;
        }
        if (c.calls++ == 0 || failed) return eager();
        if (!c.exec) {
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                failed = true;
                (void)hipGetLastError();
                return eager();
            }
            const int err = eager();
            hipGraph_t graph = nullptr;
            const hipError_t e_end = hipStreamEndCapture(s, &graph);
            if (err || e_end != hipSuccess || !graph ||
                hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                c.exec = nullptr;
                failed = true;
                (void)hipGetLastError();
                return eager();  // nothing ran during the capture: run it now
            }
            (void)hipGraphDestroy(graph);
            c.cx = x;
            c.cy = y;
            c.cs = s;
            c.epoch = now;
        }
        if (x == c.cx && y == c.cy && s == c.cs) {
            replays++;
            SFEM_AMD_HIP_CHECK(hipGraphLaunch(c.exec, s));
            return SFEM_SUCCESS;
        }
        return eager();
    }
};

// Runs one of two operators of the same shape; the CG fallback hook selects the second one
// when the primary (e.g. a learned smoother) stops reducing the residual.
struct SwitchOp final : OpBase {
    std::shared_ptr<OpBase> ops[2];
    int active{0};
    int events{0};
    int real_type() const override { return ops[0]->real_type(); }
    ptrdiff_t rows() const override { return ops[0]->rows(); }
    int apply(const void *x, void *y, hipStream_t s) override { return ops[active]->apply(x, y, s); }
    void set_initial_guess_zero(bool v) override {
        ops[0]->set_initial_guess_zero(v);
        ops[1]->set_initial_guess_zero(v);
    }
    bool zero_guess_ignores_x() const override {
        return ops[0]->zero_guess_ignores_x() && ops[1]->zero_guess_ignores_x();
    }
};

// zf = P32(rf) from zf = 0 on a precision bridge's own FP32 buffers (rf: its xf, zf: its yf): the
// FP32 part of the bridge's overwriting apply, without the casts around it
struct BridgeCore final : OpBase {
    std::shared_ptr<PrecisionBridge> bridge;
    int real_type() const override { return SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return bridge->rows(); }
    int apply(const void *x, void *y, hipStream_t s) override { return apply_overwrite(x, y, s); }
    int apply_overwrite(const void *x, void *y, hipStream_t s) override {
        if (blas_values<float>(rows(), 0.f, (float *)y, s)) return SFEM_FAILURE;
        bridge->op32->set_initial_guess_zero(true);
        const int err = bridge->op32->apply(x, y, s);
        bridge->op32->set_initial_guess_zero(false);
        return err;
    }
};

// Flexible PCG (the steps of sfem::ConjugateGradient<double>::flexible_pcg) for an FP32
// preconditioner behind a precision bridge, with the preconditioned residual kept in FP32: the CG
// update also writes the FP32 residual the preconditioner reads, and the dots and the direction
// update read z in FP32, so neither of the bridge's casts nor an FP64 copy of z is made (32 of the
// 140 bytes per unknown that an iteration's vector operations move).  Each step is the FP64 step
// on (double) z, which the bridge's cast makes exactly: the iterates are the same.
struct MixedFlexibleCG final : sfem::MatrixFreeLinearSolver<double> {
    std::shared_ptr<OpBase> A;                // FP64 operator
    std::shared_ptr<PrecisionBridge> bridge;  // its xf and yf hold rf and zf
    std::shared_ptr<OpBase> core;             // BridgeCore, replayed as a graph when the CG's was
    double rtol{1e-10}, atol{1e-16};
    int max_it{10000};
    int residual_replacement{50};
    int iterations_{0};
    int last_status{SFEM_FAILURE};
    std::vector<double> history;
    std::function<void(int, double, double)> on_iteration;
    DevArray<double> r_, p_, Ap_;

    int iterations() const override { return iterations_; }
    void set_op(const sfem::SharedOperator<double> &) override {}
    void set_preconditioner_op(const sfem::SharedOperator<double> &) override {}
    void set_max_it(int it) override { max_it = it; }
    void set_n_dofs(std::ptrdiff_t) override {}
    std::ptrdiff_t rows() const override { return A->rows(); }
    std::ptrdiff_t cols() const override { return A->rows(); }
    sfem::ExecutionSpace execution_space() const override { return sfem::EXECUTION_SPACE_DEVICE; }

    int apply(const double *const b, double *const x) override {
        history.clear();
        last_status = solve(b, x);
        return last_status;
    }

private:
    int residual(const double *b, const double *x, double *r, hipStream_t s) {
        if (A->apply_overwrite(x, r, s)) return SFEM_FAILURE;
        return blas_axpby<double>(rows(), 1.0, b, -1.0, r, s);
    }
    double dot(const double *x, const double *y, hipStream_t s) {
        double v = NAN;
        return blas_dot<double>(rows(), x, y, reduction_ws(), &v, s) ? NAN : v;
    }
    double dot(const double *x, const float *y, hipStream_t s) {
        double v = NAN;
        return blas_dot_mixed(rows(), x, y, reduction_ws(), &v, s) ? NAN : v;
    }
    // rf = (float) r was written; zf = P32(rf)
    int precondition(hipStream_t s) { return core->apply_overwrite(bridge->xf.p, bridge->yf.p, s); }

    int solve(const double *b, double *x) {
        const ptrdiff_t n = rows();
        hipStream_t s = as_stream(nullptr);
        if ((ptrdiff_t)r_.n != n && (r_.alloc(n) || p_.alloc(n) || Ap_.alloc(n))) return SFEM_FAILURE;
        if ((ptrdiff_t)bridge->xf.n != n && (bridge->xf.alloc(n) || bridge->yf.alloc(n))) return SFEM_FAILURE;
        if (!reduction_ws()) return SFEM_FAILURE;
        double *r = r_.p, *p = p_.p, *Ap = Ap_.p;
        float *rf = bridge->xf.p, *zf = bridge->yf.p;
        const double bnorm = std::sqrt(dot(b, b, s));
        if (residual(b, x, r, s)) return SFEM_FAILURE;
        double rnorm = std::sqrt(dot(r, r, s));
        iterations_ = 0;
        const double ref = bnorm > 0 ? bnorm : 1.0;
        if (rnorm == 0 || rnorm / ref < rtol || rnorm < atol) return SFEM_SUCCESS;
        if (blas_cast<double, float>(n, r, rf, s) || precondition(s)) return SFEM_FAILURE;
        if (blas_cast<float, double>(n, zf, p, s)) return SFEM_FAILURE;
        double rtz = dot(r, zf, s);
        for (iterations_ = 0; iterations_ < max_it; iterations_++) {
            if (A->apply_overwrite(p, Ap, s)) return SFEM_FAILURE;
            const double ptAp = dot(p, Ap, s);
            const double alpha = rtz / ptAp;
            const bool replace = residual_replacement > 0 && (iterations_ + 1) % residual_replacement == 0;
            double rtr = 0, rtzold = 0;
            if (cg_update_mixed(n, alpha, p, Ap, x, r, rf, replace ? nullptr : zf, reduction_ws(), &rtr, &rtzold, s))
                return SFEM_FAILURE;
            bool cast_r = false;  // rf no longer holds (float) r
            if (replace) {
                if (residual(b, x, r, s)) return SFEM_FAILURE;
                rtr = dot(r, r, s);
                rtzold = dot(r, zf, s);
                cast_r = true;
            }
            const double rprev = rnorm;
            rnorm = std::sqrt(rtr);
            history.push_back(rnorm / ref);
            if (on_iteration) on_iteration(iterations_, rprev, rnorm);
            if (rnorm / ref < rtol || rnorm < atol) {
                // confirm with the true residual
                if (residual(b, x, r, s)) return SFEM_FAILURE;
                const double true_norm = std::sqrt(dot(r, r, s));
                if (true_norm / ref < rtol || true_norm < atol) {
                    iterations_++;
                    if (!history.empty()) history.back() = true_norm / ref;
                    return SFEM_SUCCESS;
                }
                rnorm = true_norm;
                rtzold = dot(r, zf, s);
                cast_r = true;
            }
            if (cast_r && blas_cast<double, float>(n, r, rf, s)) return SFEM_FAILURE;
            if (precondition(s)) return SFEM_FAILURE;
            const double rtz_new = dot(r, zf, s);
            const double beta = (rtz_new - rtzold) / rtz;
            rtz = rtz_new;
            if (blas_axpby_mixed(n, 1.0, zf, beta, p, s)) return SFEM_FAILURE;
        }
        return SFEM_FAILURE;
    }
};
}  // namespace

#define CHECK_HANDLE(h)                                  \
    do {                                                 \
        if (!(h)) {                                      \
            set_last_error("null handle");               \
            return SFEM_FAILURE;                         \
        }                                                \
    } while (0)

// geometry arrays may live in host or device memory (device: transformed in place on the device)
static bool device_pointer(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // plain host memory: not an error for this query
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

extern "C" {

const char *sfem_amd_last_error(void) { return g_last_error.c_str(); }

int sfem_amd_set_option(const char *name, int value) {
    SFEM_AMD_REQUIRE(name, "set_option: name");
    if (std::strcmp(name, "fused_chebyshev") == 0) {
        g_fused_chebyshev = value != 0;
        bump_state_epoch();
        return SFEM_SUCCESS;
    }
    if (std::strcmp(name, "mixed_cg") == 0) {
        g_mixed_cg = value != 0;
        return SFEM_SUCCESS;
    }
    if (std::strcmp(name, "atomic_skeleton") == 0) {
        g_atomic_skeleton = value != 0;
        bump_state_epoch();
        return SFEM_SUCCESS;
    }
    if (std::strcmp(name, "hex8_patches") == 0) {
        g_hex8_patches = value != 0;
        bump_state_epoch();
        return SFEM_SUCCESS;
    }
    if (std::strcmp(name, "elasticity_form") == 0) {
        SFEM_AMD_REQUIRE(value >= 0 && value <= 2, "set_option: elasticity_form is 0 (auto), 1 (element) or 2 (stencil)");
        g_elasticity_form = value;
        bump_state_epoch();
        return SFEM_SUCCESS;
    }
    set_last_error(std::string("set_option: unknown option ") + name);
    return SFEM_FAILURE;
}

int sfem_amd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sfem_amd_device_synchronize(void) {
    SFEM_AMD_HIP_CHECK(hipDeviceSynchronize());
    return SFEM_SUCCESS;
}

void *sfem_amd_malloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        set_last_error("hipMalloc failed");
        return nullptr;
    }
    return p;
}
int sfem_amd_free(void *p) {
    SFEM_AMD_HIP_CHECK(hipFree(p));
    return SFEM_SUCCESS;
}
int sfem_amd_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    SFEM_AMD_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SFEM_SUCCESS;
}
int sfem_amd_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    SFEM_AMD_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SFEM_SUCCESS;
}
int sfem_amd_memset(void *dst, int v, size_t bytes, void *stream) {
    if (!bytes) return SFEM_SUCCESS;
    SFEM_AMD_HIP_CHECK(hipMemsetAsync(dst, v, bytes, as_stream(stream)));
    return SFEM_SUCCESS;
}

// ----------------------------------------------------------------- meshes
static int nodes_per_element(int type, int level) {
    switch (type) {
        case SFEM_AMD_TET4:
            return 4;
        case SFEM_AMD_HEX8:
            return 8;
        case SFEM_AMD_SSHEX8:
            return nxe_of(level);
        default:
            return -1;
    }
}

int sfem_amd_mesh_create(int type, int level, ptrdiff_t ne, ptrdiff_t stride, const int32_t *el, ptrdiff_t nnodes,
                         sfem_amd_mesh **out) {
    const int nv = nodes_per_element(type, level);
    SFEM_AMD_REQUIRE(nv > 0, "mesh_create: unknown element type");
    SFEM_AMD_REQUIRE(type != SFEM_AMD_SSHEX8 || sshex8_level_supported(level), "mesh_create: unsupported sshex8 level");
    SFEM_AMD_REQUIRE(stride >= ne, "mesh_create: stride < nelements");
    for (ptrdiff_t i = 0; i < (ptrdiff_t)nv; i++)
        for (ptrdiff_t e = 0; e < ne; e++) {
            const int32_t g = el[i * stride + e];
            SFEM_AMD_REQUIRE(g >= 0 && g < nnodes, "mesh_create: node id out of range");
        }
    auto m = std::make_unique<sfem_amd_mesh>();
    m->type = type;
    m->level = type == SFEM_AMD_SSHEX8 ? level : 1;
    m->E = ne;
    m->nnodes = nnodes;
    m->nv = nv;
    std::vector<int32_t> h((size_t)nv * ne);
    if (type == SFEM_AMD_SSHEX8) {
        for (ptrdiff_t e = 0; e < ne; e++)
            for (int v = 0; v < nv; v++) h[(size_t)e * nv + v] = el[v * stride + e];
    } else {
        for (int v = 0; v < nv; v++) std::memcpy(&h[(size_t)v * ne], &el[v * stride], sizeof(int32_t) * ne);
    }
    if (m->idx.upload(h.data(), h.size())) return SFEM_FAILURE;
    *out = m.release();
    return SFEM_SUCCESS;
}

int sfem_amd_sshex8_refine_device(int level, ptrdiff_t nelements, ptrdiff_t nnodes, const int32_t *macro_elements,
                                  sfem_amd_mesh **out, void *stream) {
    SFEM_AMD_REQUIRE(out, "sshex8_refine_device: output handle");
    SFEM_AMD_REQUIRE(sshex8_level_supported(level), "sshex8_refine_device: unsupported sshex8 level");
    SFEM_AMD_REQUIRE(nelements == 0 || (macro_elements && device_pointer(macro_elements)),
                     "sshex8_refine_device: macro elements must be device memory");
    auto m = std::make_unique<sfem_amd_mesh>();
    m->type = SFEM_AMD_SSHEX8;
    m->level = level;
    m->E = nelements;
    m->nv = nxe_of(level);
    ptrdiff_t nt = 0;
    if (sshex8_refine_device(level, nelements, nnodes, macro_elements, m->idx, nt, as_stream(stream)))
        return SFEM_FAILURE;
    m->nnodes = nt;
    *out = m.release();
    return SFEM_SUCCESS;
}

int sfem_amd_sshex8_refine_points_device(const sfem_amd_mesh *fine, ptrdiff_t nnodes_macro,
                                         const int32_t *macro_elements, const float *macro_points, float *points,
                                         void *stream) {
    CHECK_HANDLE(fine);
    SFEM_AMD_REQUIRE(fine->type == SFEM_AMD_SSHEX8, "sshex8_refine_points_device: sshex8 mesh required");
    SFEM_AMD_REQUIRE(fine->E == 0 || (device_pointer(macro_elements) && device_pointer(macro_points) &&
                                      device_pointer(points)),
                     "sshex8_refine_points_device: device memory required");
    return sshex8_refine_points_device(fine->level, fine->E, nnodes_macro, fine->nnodes, macro_elements,
                                       macro_points, fine->idx.p, points, as_stream(stream));
}

int sfem_amd_mesh_fff(const sfem_amd_mesh *m, const float *points, float *fff, void *stream) {
    CHECK_HANDLE(m);
    SFEM_AMD_REQUIRE(m->E == 0 || (device_pointer(points) && device_pointer(fff)), "mesh_fff: device memory required");
    return mesh_geometry_device(m->type, m->level, m->E, m->idx.p, m->nnodes, points, 0, fff, nullptr,
                                as_stream(stream));
}

int sfem_amd_mesh_adjugate_det(const sfem_amd_mesh *m, const float *points, float *adj, float *det, void *stream) {
    CHECK_HANDLE(m);
    SFEM_AMD_REQUIRE(m->type != SFEM_AMD_TET4, "mesh_adjugate_det: HEX8 or SSHEX8");
    SFEM_AMD_REQUIRE(m->E == 0 || (device_pointer(points) && device_pointer(adj) && device_pointer(det)),
                     "mesh_adjugate_det: device memory required");
    return mesh_geometry_device(m->type, m->level, m->E, m->idx.p, m->nnodes, points, 1, adj, det, as_stream(stream));
}

int sfem_amd_mesh_derefine(const sfem_amd_mesh *fine, int level, sfem_amd_mesh **out) {
    CHECK_HANDLE(fine);
    SFEM_AMD_REQUIRE(fine->type == SFEM_AMD_SSHEX8, "derefine: sshex8 mesh required");
    SFEM_AMD_REQUIRE(level >= 1 && fine->level % level == 0, "derefine: level must divide the fine level");
    auto m = std::make_unique<sfem_amd_mesh>();
    m->E = fine->E;
    if (level == 1) {
        m->type = SFEM_AMD_HEX8;
        m->level = 1;
        m->nv = 8;
        if (m->idx.alloc((size_t)8 * m->E)) return SFEM_FAILURE;
        if (sshex8_to_hex8(m->E, fine->level, fine->idx.p, m->idx.p, nullptr)) return SFEM_FAILURE;
    } else {
        SFEM_AMD_REQUIRE(sshex8_level_supported(level), "derefine: unsupported level");
        m->type = SFEM_AMD_SSHEX8;
        m->level = level;
        m->nv = nxe_of(level);
        if (m->idx.alloc((size_t)m->nv * m->E)) return SFEM_FAILURE;
        if (sshex8_derefine(m->E, fine->level, level, fine->idx.p, m->idx.p, nullptr)) return SFEM_FAILURE;
    }
    m->parent = fine;
    // nested levels are prefixes of the fine numbering: node count = max id + 1
    std::vector<int32_t> h(m->idx.n);
    SFEM_AMD_HIP_CHECK(hipMemcpy(h.data(), m->idx.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost));
    int32_t mx = -1;
    for (int32_t v : h) mx = v > mx ? v : mx;
    m->nnodes = (ptrdiff_t)mx + 1;
    *out = m.release();
    return SFEM_SUCCESS;
}

ptrdiff_t sfem_amd_mesh_nnodes(const sfem_amd_mesh *m) { return m ? m->nnodes : -1; }
ptrdiff_t sfem_amd_mesh_nelements(const sfem_amd_mesh *m) { return m ? m->E : -1; }
int sfem_amd_mesh_level(const sfem_amd_mesh *m) { return m ? m->level : -1; }
int sfem_amd_mesh_element_type(const sfem_amd_mesh *m) { return m ? m->type : -1; }

int sfem_amd_mesh_download(const sfem_amd_mesh *m, int32_t *el) {
    CHECK_HANDLE(m);
    std::vector<int32_t> h(m->idx.n);
    SFEM_AMD_HIP_CHECK(hipMemcpy(h.data(), m->idx.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost));
    if (m->type == SFEM_AMD_SSHEX8) {
        for (ptrdiff_t e = 0; e < m->E; e++)
            for (int v = 0; v < m->nv; v++) el[(size_t)v * m->E + e] = h[(size_t)e * m->nv + v];
    } else {
        std::memcpy(el, h.data(), sizeof(int32_t) * h.size());
    }
    return SFEM_SUCCESS;
}

void sfem_amd_mesh_destroy(sfem_amd_mesh *m) { delete m; }

// ----------------------------------------------------------------- operators
static int soa_upload(DevArray<float> &dst, const float *aos, ptrdiff_t E, int k) {
    if (device_pointer(aos)) {
        if (dst.alloc((size_t)k * E) || aos_to_soa_device(aos, E, k, dst.p, nullptr)) return SFEM_FAILURE;
        SFEM_AMD_HIP_CHECK(hipDeviceSynchronize());
        return SFEM_SUCCESS;
    }
    std::vector<float> h((size_t)k * E);
    for (ptrdiff_t e = 0; e < E; e++)
        for (int q = 0; q < k; q++) h[(size_t)q * E + e] = aos[e * k + q];
    return dst.upload(h.data(), h.size());
}

int sfem_amd_laplacian_create(const sfem_amd_mesh *mesh, const float *fff, const float *coeff, int rt,
                              sfem_amd_op **out) {
    CHECK_HANDLE(mesh);
    SFEM_AMD_REQUIRE(fff, "laplacian_create: fff required");
    const ptrdiff_t E = mesh->E;
    if (rt == SFEM_AMD_FLOAT64) {
        auto op = std::make_shared<LaplacianOp<double>>();
        op->m = mesh;
        SFEM_AMD_REQUIRE(mesh->type != SFEM_AMD_TET4 || !coeff, "tet4 Laplacian: no coefficient");
        if (soa_upload(op->fff32, fff, E, 6)) return SFEM_FAILURE;
        if (coeff && op->coeff.upload(coeff, E)) return SFEM_FAILURE;
        *out = wrap(op);
        return SFEM_SUCCESS;
    }
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT32, "laplacian_create: real type");
    SFEM_AMD_REQUIRE(mesh->type == SFEM_AMD_SSHEX8, "FP32 Laplacian: sshex8 meshes only");
    auto op = std::make_shared<LaplacianOp<float>>();
    op->m = mesh;
    if (device_pointer(fff)) {
        SFEM_AMD_REQUIRE(!coeff || device_pointer(coeff), "laplacian_create: fff and coeff in the same memory");
        if (op->fff16.alloc((size_t)6 * E) || aos_to_bf16_soa_device(fff, coeff, E, 6, op->fff16.p, nullptr))
            return SFEM_FAILURE;
        SFEM_AMD_HIP_CHECK(hipDeviceSynchronize());
        *out = wrap(op);
        return SFEM_SUCCESS;
    }
    std::vector<uint16_t> h((size_t)6 * E);
    for (ptrdiff_t e = 0; e < E; e++) {
        const double c = coeff ? (double)coeff[e] : 1.0;
        for (int q = 0; q < 6; q++) h[(size_t)q * E + e] = to_bf16((float)((double)fff[e * 6 + q] * c));
    }
    if (op->fff16.upload(h.data(), h.size())) return SFEM_FAILURE;
    *out = wrap(op);
    return SFEM_SUCCESS;
}

int sfem_amd_element_matrix_op_create(const sfem_amd_mesh *mesh, const double *matrices, int bs, int rt,
                                      sfem_amd_op **out) {
    CHECK_HANDLE(mesh);
    SFEM_AMD_REQUIRE(matrices && out, "element_matrix_op_create: arguments");
    SFEM_AMD_REQUIRE(mesh->type == SFEM_AMD_SSHEX8, "element_matrix_op_create: sshex8 meshes only");
    SFEM_AMD_REQUIRE(bs == 1 || bs == 3, "element_matrix_op_create: block size 1 (8x8) or 3 (24x24)");
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT64 || rt == SFEM_AMD_FLOAT32, "element_matrix_op_create: real type");
    const ptrdiff_t E = mesh->E;
    auto finish = [&](auto op, DevArray<double> &dst, size_t per) -> int {
        op->m = mesh;
        if (dst.upload(matrices, (size_t)E * per)) return SFEM_FAILURE;
        *out = wrap(op);
        return SFEM_SUCCESS;
    };
    if (bs == 1) {
        if (rt == SFEM_AMD_FLOAT64) {
            auto op = std::make_shared<LaplacianOp<double>>();
            return finish(op, op->Kem, 64);
        }
        auto op = std::make_shared<LaplacianOp<float>>();
        return finish(op, op->Kem, 64);
    }
    if (rt == SFEM_AMD_FLOAT64) {
        auto op = std::make_shared<ElasticityOp<double>>();
        return finish(op, op->K, 576);
    }
    auto op = std::make_shared<ElasticityOp<float>>();
    return finish(op, op->K, 576);
}

int sfem_amd_linear_elasticity_create(const sfem_amd_mesh *mesh, const float *adj, const float *det, double mu,
                                      double lambda, const float *coeff, int rt, sfem_amd_op **out) {
    CHECK_HANDLE(mesh);
    SFEM_AMD_REQUIRE(adj && det, "elasticity_create: adjugate and determinant required");
    SFEM_AMD_REQUIRE(mesh->type == SFEM_AMD_HEX8 || mesh->type == SFEM_AMD_SSHEX8, "elasticity: HEX8 or SSHEX8");
    const ptrdiff_t E = mesh->E;
    auto setup = [&](auto op) -> int {
        op->m = mesh;
        op->mu = mu;
        op->lambda = lambda;
        if (mesh->type == SFEM_AMD_HEX8) {
            if (coeff && op->coeff.upload(coeff, E)) return SFEM_FAILURE;
            if (rt == SFEM_AMD_FLOAT64)
                return (soa_upload(op->adj, adj, E, 9) || op->det.upload(det, E)) ? SFEM_FAILURE : SFEM_SUCCESS;
            // FP32 operator: bf16 SoA words
            if (device_pointer(adj)) {
                SFEM_AMD_REQUIRE(device_pointer(det), "elasticity_create: adjugate and det in the same memory");
                if (op->adj16.alloc((size_t)9 * E) || op->det16.alloc((size_t)E) ||
                    aos_to_bf16_soa_device(adj, nullptr, E, 9, op->adj16.p, nullptr) ||
                    aos_to_bf16_soa_device(det, nullptr, E, 1, op->det16.p, nullptr))
                    return SFEM_FAILURE;
                SFEM_AMD_HIP_CHECK(hipDeviceSynchronize());
                return SFEM_SUCCESS;
            }
            std::vector<uint16_t> ha((size_t)9 * E), hd(E);
            for (ptrdiff_t e = 0; e < E; e++) {
                for (int q = 0; q < 9; q++) ha[(size_t)q * E + e] = to_bf16(adj[e * 9 + q]);
                hd[e] = to_bf16(det[e]);
            }
            return (op->adj16.upload(ha.data(), ha.size()) || op->det16.upload(hd.data(), hd.size())) ? SFEM_FAILURE
                                                                                                     : SFEM_SUCCESS;
        }
        // sshex8: per-macro 24x24 matrices, built on device from the float32 words (FP64
        // path) or from BF16-rounded words (FP32 path)
        DevArray<float> dcoeff;
        if (coeff && dcoeff.upload(coeff, E)) return SFEM_FAILURE;
        if (op->K.alloc((size_t)E * 576)) return SFEM_FAILURE;
        if (rt == SFEM_AMD_FLOAT64) {
            DevArray<float> a, d;
            if (soa_upload(a, adj, E, 9) || d.upload(det, E)) return SFEM_FAILURE;
            if (sshex8_elasticity_matrices<float>(mesh->level, E, a.p, d.p, E, mu, lambda, dcoeff.p, op->K.p, nullptr))
                return SFEM_FAILURE;
        } else {
            DevArray<uint16_t> a, d;
            if (device_pointer(adj)) {
                SFEM_AMD_REQUIRE(device_pointer(det), "elasticity_create: adjugate and det in the same memory");
                if (a.alloc((size_t)9 * E) || d.alloc((size_t)E) ||
                    aos_to_bf16_soa_device(adj, nullptr, E, 9, a.p, nullptr) ||
                    aos_to_bf16_soa_device(det, nullptr, E, 1, d.p, nullptr))
                    return SFEM_FAILURE;
            } else {
                std::vector<uint16_t> ha((size_t)9 * E), hd(E);
                for (ptrdiff_t e = 0; e < E; e++) {
                    for (int q = 0; q < 9; q++) ha[(size_t)q * E + e] = to_bf16(adj[e * 9 + q]);
                    hd[e] = to_bf16(det[e]);
                }
                if (a.upload(ha.data(), ha.size()) || d.upload(hd.data(), hd.size())) return SFEM_FAILURE;
            }
            if (sshex8_elasticity_matrices<uint16_t>(mesh->level, E, a.p, d.p, E, mu, lambda, dcoeff.p, op->K.p,
                                                     nullptr))
                return SFEM_FAILURE;
        }
        SFEM_AMD_HIP_CHECK(hipDeviceSynchronize());
        return SFEM_SUCCESS;
    };
    if (rt == SFEM_AMD_FLOAT64) {
        auto op = std::make_shared<ElasticityOp<double>>();
        if (setup(op)) return SFEM_FAILURE;
        *out = wrap(op);
    } else {
        auto op = std::make_shared<ElasticityOp<float>>();
        if (setup(op)) return SFEM_FAILURE;
        op->matrix_cores = true;  // sshex8: its BF16 geometry already bounds the accuracy (~1e-2)
        *out = wrap(op);
    }
    return SFEM_SUCCESS;
}

// host-side set-up inside an apply (maps, staging vectors) may throw: no exception crosses the C ABI
static int guarded(const std::function<int()> &f) {
    try {
        return f();
    } catch (const std::exception &e) {
        set_last_error(std::string("host exception: ") + e.what());
    } catch (...) {
        set_last_error("host exception");
    }
    return SFEM_FAILURE;
}
int sfem_amd_op_apply(sfem_amd_op *op, const void *x, void *y, void *stream) {
    CHECK_HANDLE(op);
    return guarded([&] { return op->impl->apply(x, y, as_stream(stream)); });
}
int sfem_amd_op_diag(sfem_amd_op *op, void *d, void *stream) {
    CHECK_HANDLE(op);
    return guarded([&] { return op->impl->diag(d, as_stream(stream)); });
}
int sfem_amd_op_block_diag_sym(sfem_amd_op *op, void *d, void *stream) {
    CHECK_HANDLE(op);
    return guarded([&] { return op->impl->block_diag(d, as_stream(stream)); });
}
ptrdiff_t sfem_amd_op_rows(const sfem_amd_op *op) { return op ? op->impl->rows() : -1; }
ptrdiff_t sfem_amd_op_cols(const sfem_amd_op *op) { return op ? op->impl->cols() : -1; }
int sfem_amd_op_real_type(const sfem_amd_op *op) { return op ? op->impl->real_type() : -1; }
void sfem_amd_op_destroy(sfem_amd_op *op) { delete op; }
int sfem_amd_op_iterations(const sfem_amd_op *op) { return op ? op->impl->iterations() : -1; }
int sfem_amd_op_last_status(const sfem_amd_op *op) { return op ? op->impl->last_status() : SFEM_FAILURE; }
int sfem_amd_cg_history(const sfem_amd_op *op, double *out, int cap) {
    if (!op) return -1;
    const std::vector<double> *h = op->impl->history();
    if (!h) return 0;
    const int n = (int)h->size() < cap ? (int)h->size() : cap;
    for (int i = 0; i < n; i++) out[i] = (*h)[i];
    return (int)h->size();
}

int sfem_amd_constrained_op_create(sfem_amd_op *op, ptrdiff_t n_nodes, const int32_t *nodes, int bs, int nc,
                                   const int *comps, sfem_amd_op **out) {
    CHECK_HANDLE(op);
    SFEM_AMD_REQUIRE(bs >= 1, "constrained_op: block size");
    std::vector<int> cs;
    if (comps)
        cs.assign(comps, comps + nc);
    else
        for (int c = 0; c < bs; c++) cs.push_back(c);
    std::vector<int32_t> dofs;
    dofs.reserve((size_t)n_nodes * cs.size());
    for (int c : cs) {
        SFEM_AMD_REQUIRE(c >= 0 && c < bs, "constrained_op: component out of range");
        for (ptrdiff_t i = 0; i < n_nodes; i++) {
            const int64_t d = (int64_t)nodes[i] * bs + c;
            SFEM_AMD_REQUIRE(d >= 0 && d < op->impl->rows(), "constrained_op: node out of range");
            dofs.push_back((int32_t)d);
        }
    }
    auto c = std::make_shared<ConstrainedOp>();
    c->inner = op->impl;
    if (c->dofs.upload(dofs.data(), dofs.size())) return SFEM_FAILURE;
    *out = wrap(c);
    return SFEM_SUCCESS;
}

int sfem_amd_constrained_op_set_value(sfem_amd_op *constrained, double value, void *x, void *stream) {
    CHECK_HANDLE(constrained);
    auto *c = dynamic_cast<ConstrainedOp *>(constrained->impl.get());
    SFEM_AMD_REQUIRE(c, "not a constrained operator");
    return c->set_value(value, x, as_stream(stream));
}

// ----------------------------------------------------------------- transfers
static int transfer_create(const sfem_amd_mesh *fine, const sfem_amd_mesh *coarse, int bs, int rt,
                           const sfem_amd_op *constrained, bool prolong, sfem_amd_op **out) {
    CHECK_HANDLE(fine);
    CHECK_HANDLE(coarse);
    SFEM_AMD_REQUIRE(fine->type == SFEM_AMD_SSHEX8, "transfer: fine mesh must be sshex8");
    SFEM_AMD_REQUIRE(fine->E == coarse->E, "transfer: meshes must share macro elements");
    SFEM_AMD_REQUIRE(fine->level % coarse->level == 0 && coarse->level < fine->level, "transfer: nested levels");
    SFEM_AMD_REQUIRE(!constrained || dynamic_cast<ConstrainedOp *>(constrained->impl.get()),
                     "transfer: constraints must be a constrained operator");
    if (coarse->parent != fine) {
        // The kernels address coarse dofs through the fine connectivity (nested numbering):
        // every coarse-position id must index the coarse vector.
        DevArray<int32_t> tmp;
        if (tmp.alloc((size_t)nxe_of(coarse->level) * fine->E)) return SFEM_FAILURE;
        if (sshex8_derefine(fine->E, fine->level, coarse->level, fine->idx.p, tmp.p, nullptr)) return SFEM_FAILURE;
        std::vector<int32_t> h(tmp.n);
        SFEM_AMD_HIP_CHECK(hipMemcpy(h.data(), tmp.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost));
        for (int32_t v : h) SFEM_AMD_REQUIRE(v >= 0 && v < coarse->nnodes, "transfer: meshes are not nested (coarse id out of range)");
    }
    auto make = [&](auto op) {
        op->fine = fine;
        op->Lc = coarse->level;
        op->n_coarse = coarse->nnodes;
        op->vs = bs;
        op->prolong = prolong;
        if (constrained) op->constrained = constrained->impl;
        *out = wrap(op);
    };
    if (rt == SFEM_AMD_FLOAT64)
        make(std::make_shared<TransferOp<double>>());
    else
        make(std::make_shared<TransferOp<float>>());
    // build the lazily needed tables now (setup time)
    if (prolong ? mesh_owner(fine, nullptr) : mesh_counts(fine, nullptr)) return SFEM_FAILURE;
    return SFEM_SUCCESS;
}

int sfem_amd_prolongation_create(const sfem_amd_mesh *coarse, const sfem_amd_mesh *fine, int bs, int rt,
                                 const sfem_amd_op *constrained, sfem_amd_op **out) {
    return transfer_create(fine, coarse, bs, rt, constrained, true, out);
}
int sfem_amd_restriction_create(const sfem_amd_mesh *fine, const sfem_amd_mesh *coarse, int bs, int rt,
                                const sfem_amd_op *constrained, sfem_amd_op **out) {
    return transfer_create(fine, coarse, bs, rt, constrained, false, out);
}

// ----------------------------------------------------------------- BLAS
#define DISPATCH_RT(rt, CALL_D, CALL_F)                               \
    do {                                                              \
        if ((rt) == SFEM_AMD_FLOAT64) return CALL_D;                  \
        if ((rt) == SFEM_AMD_FLOAT32) return CALL_F;                  \
        set_last_error("unknown real type");                          \
        return SFEM_FAILURE;                                          \
    } while (0)

int sfem_amd_blas_axpby(int rt, ptrdiff_t n, double a, const void *x, double b, void *y, void *s) {
    DISPATCH_RT(rt, blas_axpby<double>(n, a, (const double *)x, b, (double *)y, as_stream(s)),
                blas_axpby<float>(n, (float)a, (const float *)x, (float)b, (float *)y, as_stream(s)));
}
int sfem_amd_blas_zaxpby(int rt, ptrdiff_t n, double a, const void *x, double b, const void *y, void *z, void *s) {
    DISPATCH_RT(rt, blas_zaxpby<double>(n, a, (const double *)x, b, (const double *)y, (double *)z, as_stream(s)),
                blas_zaxpby<float>(n, (float)a, (const float *)x, (float)b, (const float *)y, (float *)z, as_stream(s)));
}
int sfem_amd_blas_scal(int rt, ptrdiff_t n, double a, void *x, void *s) {
    DISPATCH_RT(rt, blas_scal<double>(n, a, (double *)x, as_stream(s)), blas_scal<float>(n, (float)a, (float *)x, as_stream(s)));
}
int sfem_amd_blas_values(int rt, ptrdiff_t n, double v, void *x, void *s) {
    DISPATCH_RT(rt, blas_values<double>(n, v, (double *)x, as_stream(s)), blas_values<float>(n, (float)v, (float *)x, as_stream(s)));
}
int sfem_amd_blas_copy(int rt, ptrdiff_t n, const void *x, void *y, void *s) {
    DISPATCH_RT(rt, blas_copy<double>(n, (const double *)x, (double *)y, as_stream(s)),
                blas_copy<float>(n, (const float *)x, (float *)y, as_stream(s)));
}
int sfem_amd_blas_reciprocal(int rt, ptrdiff_t n, double a, void *x, void *s) {
    DISPATCH_RT(rt, blas_reciprocal<double>(n, a, (double *)x, as_stream(s)),
                blas_reciprocal<float>(n, (float)a, (float *)x, as_stream(s)));
}
int sfem_amd_blas_xypaz(int rt, ptrdiff_t n, const void *x, const void *y, double a, void *z, void *s) {
    DISPATCH_RT(rt, blas_xypaz<double>(n, (const double *)x, (const double *)y, a, (double *)z, as_stream(s)),
                blas_xypaz<float>(n, (const float *)x, (const float *)y, (float)a, (float *)z, as_stream(s)));
}
double sfem_amd_blas_dot(int rt, ptrdiff_t n, const void *x, const void *y, void *s) {
    double r = NAN;
    double *ws = reduction_ws();
    if (!ws) return NAN;
    int err = rt == SFEM_AMD_FLOAT64 ? blas_dot<double>(n, (const double *)x, (const double *)y, ws, &r, as_stream(s))
                                     : blas_dot<float>(n, (const float *)x, (const float *)y, ws, &r, as_stream(s));
    return err ? NAN : r;
}
double sfem_amd_blas_norm2(int rt, ptrdiff_t n, const void *x, void *s) { return std::sqrt(sfem_amd_blas_dot(rt, n, x, x, s)); }
int sfem_amd_blas_dot_noalloc(int rt, ptrdiff_t n, const void *x, const void *y, double *result, void *s) {
    DISPATCH_RT(rt, blas_dot_noalloc<double>(n, (const double *)x, (const double *)y, result, as_stream(s)),
                blas_dot_noalloc<float>(n, (const float *)x, (const float *)y, result, as_stream(s)));
}
int sfem_amd_blas_cast(int rt_from, ptrdiff_t n, const void *x, int rt_to, void *y, void *s) {
    if (rt_from == SFEM_AMD_FLOAT64 && rt_to == SFEM_AMD_FLOAT32) return blas_cast<double, float>(n, (const double *)x, (float *)y, as_stream(s));
    if (rt_from == SFEM_AMD_FLOAT32 && rt_to == SFEM_AMD_FLOAT64) return blas_cast<float, double>(n, (const float *)x, (double *)y, as_stream(s));
    if (rt_from == rt_to) return sfem_amd_blas_copy(rt_from, n, x, y, s);
    set_last_error("cast: unknown real types");
    return SFEM_FAILURE;
}
double sfem_amd_cg_update_x_r(int rt, ptrdiff_t n, double a, const void *p, const void *Ap, void *x, void *r,
                              int with_rtr, void *s) {
    double *ws = reduction_ws();
    double rtr = 0;
    int err = rt == SFEM_AMD_FLOAT64
                      ? cg_update<double>(n, a, (const double *)p, (const double *)Ap, (double *)x, (double *)r, ws,
                                          with_rtr ? &rtr : nullptr, as_stream(s))
                      : cg_update<float>(n, (float)a, (const float *)p, (const float *)Ap, (float *)x, (float *)r, ws,
                                         with_rtr ? &rtr : nullptr, as_stream(s));
    return err ? NAN : rtr;
}
double sfem_amd_cg_update_x_r_dot(int rt, ptrdiff_t n, double a, const void *p, const void *Ap, void *x, void *r,
                                  const void *z, double *rz, void *s) {
    double *ws = reduction_ws();
    double rtr = 0;
    if (!ws || !z || !rz) {
        set_last_error("cg_update_x_r_dot: arguments");
        return NAN;
    }
    int err = rt == SFEM_AMD_FLOAT64
                      ? cg_update<double>(n, a, (const double *)p, (const double *)Ap, (double *)x, (double *)r, ws,
                                          &rtr, as_stream(s), (const double *)z, rz)
                      : cg_update<float>(n, (float)a, (const float *)p, (const float *)Ap, (float *)x, (float *)r, ws,
                                         &rtr, as_stream(s), (const float *)z, rz);
    return err ? NAN : rtr;
}
int sfem_amd_cheb3_update(int rt, ptrdiff_t n, double alpha, double beta, const void *rhs, const void *Ax, void *p,
                          void *x, void *s) {
    DISPATCH_RT(rt,
                cheb3_update<double>(n, alpha, beta, (const double *)rhs, (const double *)Ax, (double *)p, (double *)x,
                                     as_stream(s)),
                cheb3_update<float>(n, (float)alpha, (float)beta, (const float *)rhs, (const float *)Ax, (float *)p,
                                    (float *)x, as_stream(s)));
}

int sfem_amd_cheb3_update_diag(int rt, ptrdiff_t n, double alpha, double beta, const void *rhs, const void *Ax,
                               const void *inv_diag, void *p, void *x, void *s) {
    if (rt == SFEM_AMD_FLOAT64)
        return cheb3_update_diag<double>(n, alpha, beta, (const double *)rhs, (const double *)Ax,
                                         (const double *)inv_diag, (double *)p, (double *)x, as_stream(s));
    return cheb3_update_diag<float>(n, (float)alpha, (float)beta, (const float *)rhs, (const float *)Ax,
                                    (const float *)inv_diag, (float *)p, (float *)x, as_stream(s));
}

int sfem_amd_cheb3_start(int rt, ptrdiff_t n, double alpha, const void *rhs, const void *inv_diag, int have_ax,
                         void *p, void *x, void *s) {
    if (rt == SFEM_AMD_FLOAT64)
        return cheb3_start<double>(n, alpha, (const double *)rhs, (const double *)inv_diag, have_ax, (double *)p,
                                   (double *)x, as_stream(s));
    return cheb3_start<float>(n, (float)alpha, (const float *)rhs, (const float *)inv_diag, have_ax, (float *)p,
                              (float *)x, as_stream(s));
}

// ----------------------------------------------------------------- solvers
extern "C++" {
template <typename T>
static int cg_create_t(sfem_amd_op *A, sfem_amd_op *P, int flexible, double rtol, double atol, int max_it,
                       sfem_amd_op **out) {
    if constexpr (std::is_same<T, double>::value) {
        // flexible CG around an FP32 preconditioner behind a precision bridge (itself, or replayed as
        // a graph): the same iterates with z kept in FP32 (option "mixed_cg", default 1)
        if (P && flexible == 1 && g_mixed_cg) {
            std::shared_ptr<PrecisionBridge> br = std::dynamic_pointer_cast<PrecisionBridge>(P->impl);
            bool graph = false;
            if (!br) {
                if (auto g = std::dynamic_pointer_cast<GraphOp>(P->impl)) {
                    br = std::dynamic_pointer_cast<PrecisionBridge>(g->inner);
                    graph = br != nullptr;
                }
            }
            if (br) {
                auto cg = std::make_shared<MixedFlexibleCG>();
                cg->A = A->impl;
                cg->bridge = br;
                auto core = std::make_shared<BridgeCore>();
                core->bridge = br;
                if (graph) {
                    auto g = std::make_shared<GraphOp>();
                    g->inner = core;
                    cg->core = g;
                } else {
                    cg->core = core;
                }
                cg->rtol = rtol;
                cg->atol = atol;
                cg->max_it = max_it;
                auto s = std::make_shared<SolverOp<double>>();
                s->op = cg;
                s->n = A->impl->rows();
                MixedFlexibleCG *raw = cg.get();
                s->its = [raw] { return raw->iterations(); };
                s->status = [raw] { return raw->last_status; };
                s->hist = [raw] { return &raw->history; };
                s->keep.push_back(P->impl);
                *out = wrap(s);
                return SFEM_SUCCESS;
            }
        }
    }
    auto cg = std::make_shared<sfem::ConjugateGradient<T>>();
    cg->default_init();
    cg->set_op(view<T>(A->impl, true));
    cg->apply_overwrites_output = true;
    if (P) {
        // z = P r from zero through the operator's overwriting apply: its default is the zero fill
        // and the apply, operators that can write z directly (the precision bridge) skip the fill
        cg->set_preconditioner_op(view<T>(P->impl, true));
        cg->preconditioner_overwrites_output = true;
    }
    cg->flexible = flexible == 1;
    cg->stop_on_residual_norm = flexible == 2;
    cg->rtol = (T)rtol;
    cg->atol = (T)atol;
    cg->max_it = max_it;
    auto s = std::make_shared<SolverOp<T>>();
    s->op = cg;
    s->n = A->impl->rows();
    sfem::ConjugateGradient<T> *raw = cg.get();
    s->its = [raw] { return raw->iterations(); };
    s->status = [raw] { return raw->last_status; };
    s->hist = [raw] { return &raw->history; };
    *out = wrap(s);
    return SFEM_SUCCESS;
}
}  // extern "C++"

int sfem_amd_cg_create(sfem_amd_op *A, sfem_amd_op *P, int flexible, double rtol, double atol, int max_it,
                       sfem_amd_op **out) {
    CHECK_HANDLE(A);
    SFEM_AMD_REQUIRE(!P || P->impl->real_type() == A->impl->real_type(), "cg: preconditioner real type differs");
    if (A->impl->real_type() == SFEM_AMD_FLOAT64) return cg_create_t<double>(A, P, flexible, rtol, atol, max_it, out);
    return cg_create_t<float>(A, P, flexible, rtol, atol, max_it, out);
}

extern "C++" {
template <typename T>
static int cheb3_create_t(sfem_amd_op *A, int max_it, double smax, double smin, int eig_its, double eig_tol, int zero,
                          const void *init, sfem_amd_op **out, const void *inv_diag = nullptr, int lanczos_steps = 0) {
    auto c = std::make_shared<sfem::Chebyshev3<T>>();
    c->default_init();
    c->set_op(view<T>(A->impl, true));
    c->apply_overwrites_output = true;
    if (inv_diag) {
        const ptrdiff_t n = A->impl->rows();
        auto b = c->blas;
        c->inv_diag = sfem::Buffer<T>::own(n, b->allocate(n), [b](void *q) { b->destroy(q); });
        b->copy(n, (const T *)inv_diag, c->inv_diag->data());
    }
    c->max_it = max_it;
    c->scale_eig_max = (T)smax;
    c->scale_eig_min = (T)smin;
    c->eigen_solver_max_it = eig_its;
    c->eigen_solver_tol = (T)eig_tol;
    c->lanczos_steps = lanczos_steps;
    c->is_initial_guess_zero = zero != 0;
    if (init)
        c->init((const T *)init);
    else
        c->init_with_ones();
    if (!std::isfinite((double)c->eig_max) || c->eig_max <= 0) {
        set_last_error(lanczos_steps > 0 ? "cheb3: Lanczos estimate failed" : "cheb3: power method failed");
        return SFEM_FAILURE;
    }
    auto s = std::make_shared<SolverOp<T>>();
    s->op = c;
    s->n = A->impl->rows();
    sfem::Chebyshev3<T> *raw = c.get();
    s->its = [raw] { return raw->iterations(); };
    *out = wrap(s);
    return SFEM_SUCCESS;
}
}  // extern "C++"

int sfem_amd_cheb3_create(sfem_amd_op *A, int max_it, double smax, double smin, int eig_its, double eig_tol, int zero,
                          const void *init, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    if (A->impl->real_type() == SFEM_AMD_FLOAT64)
        return cheb3_create_t<double>(A, max_it, smax, smin, eig_its, eig_tol, zero, init, out);
    return cheb3_create_t<float>(A, max_it, smax, smin, eig_its, eig_tol, zero, init, out);
}

int sfem_amd_cheb3_diag_create(sfem_amd_op *A, const void *inv_diag, int max_it, double smax, double smin,
                               int eig_its, double eig_tol, int zero, const void *init, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    SFEM_AMD_REQUIRE(inv_diag, "cheb3_diag_create: inv_diag");
    if (A->impl->real_type() == SFEM_AMD_FLOAT64)
        return cheb3_create_t<double>(A, max_it, smax, smin, eig_its, eig_tol, zero, init, out, inv_diag);
    return cheb3_create_t<float>(A, max_it, smax, smin, eig_its, eig_tol, zero, init, out, inv_diag);
}

int sfem_amd_cheb3_lanczos_create(sfem_amd_op *A, const void *inv_diag, int max_it, double smax, double smin,
                                  int lanczos_steps, double tol, int zero, const void *init, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    SFEM_AMD_REQUIRE(lanczos_steps > 0, "cheb3_lanczos_create: lanczos_steps > 0 required");
    if (A->impl->real_type() == SFEM_AMD_FLOAT64)
        return cheb3_create_t<double>(A, max_it, smax, smin, 0, tol, zero, init, out, inv_diag, lanczos_steps);
    return cheb3_create_t<float>(A, max_it, smax, smin, 0, tol, zero, init, out, inv_diag, lanczos_steps);
}

int sfem_amd_cheb3_eigen_iterations(const sfem_amd_op *op) {
    if (!op) return -1;
    if (auto *s = dynamic_cast<SolverOp<double> *>(op->impl.get()))
        if (auto *c = dynamic_cast<sfem::Chebyshev3<double> *>(s->op.get())) return c->eigen_iterations;
    if (auto *s = dynamic_cast<SolverOp<float> *>(op->impl.get()))
        if (auto *c = dynamic_cast<sfem::Chebyshev3<float> *>(s->op.get())) return c->eigen_iterations;
    return -1;
}

double sfem_amd_cheb3_eig_max(const sfem_amd_op *op) {
    if (!op) return NAN;
    if (auto *s = dynamic_cast<SolverOp<double> *>(op->impl.get()))
        if (auto *c = dynamic_cast<sfem::Chebyshev3<double> *>(s->op.get())) return c->eig_max;
    if (auto *s = dynamic_cast<SolverOp<float> *>(op->impl.get()))
        if (auto *c = dynamic_cast<sfem::Chebyshev3<float> *>(s->op.get())) return c->eig_max;
    return NAN;
}

extern "C++" {
template <typename T>
static int jacobi_create_t(sfem_amd_op *A, const void *diag, double omega, int steps, sfem_amd_op **out) {
    auto st = std::make_shared<sfem::StationaryIteration<T>>();
    st->default_init();
    st->set_op(view<T>(A->impl));
    st->set_preconditioner_op(std::make_shared<sfem::Jacobi<T>>((const T *)diag, A->impl->rows(), (T)omega, st->blas));
    st->set_max_it(steps);
    auto s = std::make_shared<SolverOp<T>>();
    s->op = st;
    s->n = A->impl->rows();
    *out = wrap(s);
    return SFEM_SUCCESS;
}
}  // extern "C++"

extern "C++" {
// y += (omega B^-1) x with B the symmetric 3x3 diagonal blocks
template <typename T>
struct BlockInverseOp final : OpBase {
    DevArray<T> inv;
    ptrdiff_t nb{0};
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return 3 * nb; }
    int apply(const void *x, void *y, hipStream_t s) override {
        return block3_apply<T>(nb, inv.p, (const T *)x, (T *)y, s);
    }
};

template <typename T>
static int block_jacobi_create_t(sfem_amd_op *A, const void *sym, const uint8_t *mask, double omega, int steps,
                                 sfem_amd_op **out) {
    auto P = std::make_shared<BlockInverseOp<T>>();
    P->nb = A->impl->rows() / 3;
    if (P->inv.alloc((size_t)P->nb * 9) ||
        block_sym_inverse<T>(P->nb, (const T *)sym, mask, omega, P->inv.p, library_stream()))
        return SFEM_FAILURE;
    auto st = std::make_shared<sfem::StationaryIteration<T>>();
    st->default_init();
    st->set_op(view<T>(A->impl));
    st->set_preconditioner_op(view<T>(P));
    st->set_max_it(steps);
    auto s = std::make_shared<SolverOp<T>>();
    s->op = st;
    s->n = A->impl->rows();
    s->keep.push_back(P);
    *out = wrap(s);
    return SFEM_SUCCESS;
}
}  // extern "C++"

int sfem_amd_block_jacobi_create(sfem_amd_op *A, const void *block_diag_sym, const uint8_t *dof_mask, double omega,
                                 int steps, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    SFEM_AMD_REQUIRE(block_diag_sym && out && A->impl->rows() % 3 == 0, "block_jacobi_create: arguments");
    if (A->impl->real_type() == SFEM_AMD_FLOAT64)
        return block_jacobi_create_t<double>(A, block_diag_sym, dof_mask, omega, steps, out);
    return block_jacobi_create_t<float>(A, block_diag_sym, dof_mask, omega, steps, out);
}

int sfem_amd_jacobi_create(sfem_amd_op *A, const void *diag, double omega, int steps, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    if (A->impl->real_type() == SFEM_AMD_FLOAT64) return jacobi_create_t<double>(A, diag, omega, steps, out);
    return jacobi_create_t<float>(A, diag, omega, steps, out);
}

extern "C++" {
template <typename T>
static int mg_create_t(int cycle_type, int max_it, double atol, sfem_amd_op **out) {
    auto s = std::make_shared<SolverOp<T>>();
    auto mg = std::make_shared<sfem::Multigrid<T>>();
    mg->default_init();
    mg->set_cycle_type(cycle_type);
    mg->set_max_it(max_it);
    mg->set_atol((T)atol);
    s->op = mg;
    sfem::Multigrid<T> *raw = mg.get();
    s->its = [raw] { return raw->iterations(); };
    s->hist = [raw] { return &raw->residual_history; };
    *out = wrap(s);
    return SFEM_SUCCESS;
}
}  // extern "C++"

int sfem_amd_mg_create(int rt, int cycle_type, int max_it, double atol, sfem_amd_op **out) {
    if (rt == SFEM_AMD_FLOAT64) return mg_create_t<double>(cycle_type, max_it, atol, out);
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT32, "mg_create: real type");
    return mg_create_t<float>(cycle_type, max_it, atol, out);
}

extern "C++" {
template <typename T>
static int mg_add_level_t(SolverOp<T> *s, sfem_amd_op *op, sfem_amd_op *sm, sfem_amd_op *P, sfem_amd_op *R) {
    auto *mg = dynamic_cast<sfem::Multigrid<T> *>(s->op.get());
    SFEM_AMD_REQUIRE(mg, "mg_add_level: not a multigrid handle");
    mg->add_level(view<T>(op->impl, true), view<T>(sm->impl), P ? view<T>(P->impl) : nullptr,
                  R ? view<T>(R->impl, true) : nullptr);
    mg->apply_overwrites_output = true;
    mg->restriction_overwrites_output = true;  // the views call apply_overwrite
    if (s->n == 0) s->n = op->impl->rows();
    return SFEM_SUCCESS;
}
}  // extern "C++"

int sfem_amd_mg_add_level(sfem_amd_op *mgh, sfem_amd_op *op, sfem_amd_op *sm, sfem_amd_op *P, sfem_amd_op *R) {
    CHECK_HANDLE(mgh);
    CHECK_HANDLE(op);
    CHECK_HANDLE(sm);
    const int rt = op->impl->real_type();
    SFEM_AMD_REQUIRE(sm->impl->real_type() == rt && (!P || P->impl->real_type() == rt) &&
                             (!R || R->impl->real_type() == rt),
                     "mg_add_level: mixed real types");
    if (auto *s = dynamic_cast<SolverOp<double> *>(mgh->impl.get())) {
        SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT64, "mg_add_level: FP64 multigrid needs FP64 levels");
        return mg_add_level_t<double>(s, op, sm, P, R);
    }
    if (auto *s = dynamic_cast<SolverOp<float> *>(mgh->impl.get())) {
        SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT32, "mg_add_level: FP32 multigrid needs FP32 levels");
        return mg_add_level_t<float>(s, op, sm, P, R);
    }
    set_last_error("mg_add_level: not a multigrid handle");
    return SFEM_FAILURE;
}

extern "C++" {
template <typename T>
struct InverseDiagonal final : OpBase {
    DevArray<T> inv;
    int real_type() const override { return std::is_same<T, double>::value ? SFEM_AMD_FLOAT64 : SFEM_AMD_FLOAT32; }
    ptrdiff_t rows() const override { return (ptrdiff_t)inv.n; }
    int apply(const void *x, void *y, hipStream_t s) override {
        return blas_xypaz<T>((ptrdiff_t)inv.n, inv.p, (const T *)x, T(1), (T *)y, s);
    }
};
}  // extern "C++"

int sfem_amd_solver_bridge_create(sfem_amd_op *solver, sfem_amd_op **out) {
    CHECK_HANDLE(solver);
    auto b = std::make_shared<SolverBridge>();
    b->inner = solver->impl;
    *out = wrap(b);
    return SFEM_SUCCESS;
}

int sfem_amd_permutation_create(ptrdiff_t n, const int32_t *perm, int bs, int rt, int transpose,
                                const sfem_amd_op *constrained, sfem_amd_op **out) {
    SFEM_AMD_REQUIRE(perm && n >= 0 && bs >= 1, "permutation: arguments");
    std::vector<char> seen((size_t)n, 0);
    for (ptrdiff_t i = 0; i < n; i++) {
        SFEM_AMD_REQUIRE(perm[i] >= 0 && perm[i] < n && !seen[perm[i]], "permutation: not a permutation of [0, n)");
        seen[perm[i]] = 1;
    }
    SFEM_AMD_REQUIRE(!constrained || dynamic_cast<ConstrainedOp *>(constrained->impl.get()),
                     "permutation: constraints must be a constrained operator");
    auto make = [&](auto op) -> int {
        op->n = n;
        op->bs = bs;
        op->transpose = transpose != 0;
        if (constrained) op->constrained = constrained->impl;
        if (op->perm.upload(perm, (size_t)n)) return SFEM_FAILURE;
        *out = wrap(op);
        return SFEM_SUCCESS;
    };
    if (rt == SFEM_AMD_FLOAT64) return make(std::make_shared<PermutationOp<double>>());
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT32, "permutation: real type");
    return make(std::make_shared<PermutationOp<float>>());
}

int sfem_amd_chain_create(sfem_amd_op *A, sfem_amd_op *B, sfem_amd_op **out) {
    CHECK_HANDLE(A);
    CHECK_HANDLE(B);
    SFEM_AMD_REQUIRE(A->impl->real_type() == B->impl->real_type(), "chain: real types differ");
    SFEM_AMD_REQUIRE(A->impl->rows() == B->impl->cols(), "chain: sizes do not match");
    auto c = std::make_shared<ChainOp>();
    c->A = A->impl;
    c->B = B->impl;
    *out = wrap(c);
    return SFEM_SUCCESS;
}

int sfem_amd_inverse_diagonal_create(int rt, ptrdiff_t n, const void *diag, double omega, sfem_amd_op **out) {
    SFEM_AMD_REQUIRE(diag && n >= 0, "inverse_diagonal: diag required");
    auto make = [&](auto op, auto *d) -> int {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(d)>>;
        if (op->inv.alloc(n)) return SFEM_FAILURE;
        if (blas_copy<T>(n, d, op->inv.p, nullptr) || blas_reciprocal<T>(n, (T)omega, op->inv.p, nullptr))
            return SFEM_FAILURE;
        *out = wrap(op);
        return SFEM_SUCCESS;
    };
    if (rt == SFEM_AMD_FLOAT64) return make(std::make_shared<InverseDiagonal<double>>(), (const double *)diag);
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT32, "inverse_diagonal: real type");
    return make(std::make_shared<InverseDiagonal<float>>(), (const float *)diag);
}

int sfem_amd_precision_bridge_create(sfem_amd_op *op, sfem_amd_op **out) {
    CHECK_HANDLE(op);
    if (op->impl->real_type() == SFEM_AMD_FLOAT32) {
        auto b = std::make_shared<PrecisionBridge>();
        b->op32 = op->impl;
        *out = wrap(b);
        return SFEM_SUCCESS;
    }
    auto b = std::make_shared<PrecisionBridgeDown>();
    b->op64 = op->impl;
    *out = wrap(b);
    return SFEM_SUCCESS;
}

int sfem_amd_op_benchmark(sfem_amd_op *op, const void *x, void *y, int warmup, int reps, size_t flush_bytes,
                          void *stream, double *total_ms) {
    CHECK_HANDLE(op);
    SFEM_AMD_REQUIRE(reps > 0 && total_ms, "benchmark: reps > 0 and an output pointer required");
    hipStream_t s = as_stream(stream);
    for (int i = 0; i < warmup; i++)
        if (op->impl->apply(x, y, s)) return SFEM_FAILURE;
    DevArray<double> scratch;
    if (flush_bytes && scratch.alloc((flush_bytes + 7) / 8)) return SFEM_FAILURE;
    // the events are destroyed on every exit path
    struct Events {
        std::vector<hipEvent_t> v;
        ~Events() {
            for (auto &e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } events;
    std::vector<hipEvent_t> &ev = events.v;
    ev.assign(2 * (size_t)(flush_bytes ? reps : 1), nullptr);
    for (auto &e : ev) SFEM_AMD_HIP_CHECK(hipEventCreate(&e));
    int err = SFEM_SUCCESS;
    if (!flush_bytes) {
        SFEM_AMD_HIP_CHECK(hipEventRecord(ev[0], s));
        for (int i = 0; i < reps && !err; i++) err = op->impl->apply(x, y, s);
        SFEM_AMD_HIP_CHECK(hipEventRecord(ev[1], s));
    } else {
        for (int i = 0; i < reps && !err; i++) {
            err = blas_values<double>((ptrdiff_t)scratch.n, (double)i, scratch.p, s);
            SFEM_AMD_HIP_CHECK(hipEventRecord(ev[2 * i], s));
            if (!err) err = op->impl->apply(x, y, s);
            SFEM_AMD_HIP_CHECK(hipEventRecord(ev[2 * i + 1], s));
        }
    }
    SFEM_AMD_HIP_CHECK(hipStreamSynchronize(s));
    if (err) return err;  // some events were never recorded
    double total = 0;
    for (size_t i = 0; i < ev.size(); i += 2) {
        float ms = 0;
        SFEM_AMD_HIP_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        total += ms;
    }
    *total_ms = total;
    return SFEM_SUCCESS;
}

int sfem_amd_dense_op_create(int rt, ptrdiff_t n, const double *matrix, sfem_amd_op **out) {
    SFEM_AMD_REQUIRE(matrix && out && n > 0, "dense_op_create: arguments");
    SFEM_AMD_REQUIRE(rt == SFEM_AMD_FLOAT64 || rt == SFEM_AMD_FLOAT32, "dense_op_create: real type");
    auto make = [&](auto op) -> int {
        op->n = n;
        if (op->M.upload(matrix, (size_t)n * n)) return SFEM_FAILURE;
        *out = wrap(op);
        return SFEM_SUCCESS;
    };
    if (rt == SFEM_AMD_FLOAT64) return make(std::make_shared<DenseOp<double>>());
    return make(std::make_shared<DenseOp<float>>());
}

int sfem_amd_graph_op_create(sfem_amd_op *op, sfem_amd_op **out) {
    CHECK_HANDLE(op);
    SFEM_AMD_REQUIRE(out, "graph_op_create: out");
    auto g = std::make_shared<GraphOp>();
    g->inner = op->impl;
    *out = wrap(g);
    return SFEM_SUCCESS;
}

int sfem_amd_graph_op_stats(const sfem_amd_op *op, int *captured, int *replays) {
    CHECK_HANDLE(op);
    auto *g = dynamic_cast<const GraphOp *>(op->impl.get());
    SFEM_AMD_REQUIRE(g, "graph_op_stats: not a graph operator");
    if (captured) *captured = g->failed ? -1 : g->captured();
    if (replays) *replays = g->replays;
    return SFEM_SUCCESS;
}

int sfem_amd_switch_create(sfem_amd_op *primary, sfem_amd_op *fallback, sfem_amd_op **out) {
    CHECK_HANDLE(primary);
    CHECK_HANDLE(fallback);
    SFEM_AMD_REQUIRE(out, "switch_create: out");
    SFEM_AMD_REQUIRE(primary->impl->rows() == fallback->impl->rows() &&
                         primary->impl->real_type() == fallback->impl->real_type(),
                     "switch_create: operators differ in size or real type");
    auto sw = std::make_shared<SwitchOp>();
    sw->ops[0] = primary->impl;
    sw->ops[1] = fallback->impl;
    *out = wrap(sw);
    return SFEM_SUCCESS;
}

int sfem_amd_switch_select(sfem_amd_op *op, int which) {
    CHECK_HANDLE(op);
    auto *sw = dynamic_cast<SwitchOp *>(op->impl.get());
    SFEM_AMD_REQUIRE(sw && (which == 0 || which == 1), "switch_select: not a switch operator or index not 0/1");
    if (sw->active != which) bump_state_epoch();
    sw->active = which;
    return SFEM_SUCCESS;
}

int sfem_amd_switch_state(const sfem_amd_op *op, int *active, int *events) {
    CHECK_HANDLE(op);
    auto *sw = dynamic_cast<const SwitchOp *>(op->impl.get());
    SFEM_AMD_REQUIRE(sw, "switch_state: not a switch operator");
    if (active) *active = sw->active;
    if (events) *events = sw->events;
    return SFEM_SUCCESS;
}

int sfem_amd_cg_add_fallback(sfem_amd_op *cg, sfem_amd_op *sw, double max_ratio) {
    CHECK_HANDLE(cg);
    CHECK_HANDLE(sw);
    auto swp = std::dynamic_pointer_cast<SwitchOp>(sw->impl);
    SFEM_AMD_REQUIRE(swp, "cg_add_fallback: not a switch operator");
    auto hook = [swp, max_ratio](int, double prev, double now) {
        if (swp->active == 0 && prev > 0 && now > max_ratio * prev) {
            swp->active = 1;
            swp->events++;
            bump_state_epoch();
        }
    };
    auto chain = [hook](std::function<void(int, double, double)> &slot) {
        auto old = slot;
        slot = [old, hook](int it, double a, double b) {
            if (old) old(it, a, b);
            hook(it, a, b);
        };
    };
    if (auto *s = dynamic_cast<SolverOp<double> *>(cg->impl.get())) {
        if (auto *m = dynamic_cast<MixedFlexibleCG *>(s->op.get())) {
            chain(m->on_iteration);
            s->keep.push_back(swp);
            return SFEM_SUCCESS;
        }
        auto *c = dynamic_cast<sfem::ConjugateGradient<double> *>(s->op.get());
        SFEM_AMD_REQUIRE(c && c->flexible, "cg_add_fallback: needs a flexible CG (mode 1)");
        chain(c->on_iteration);
        s->keep.push_back(swp);
        return SFEM_SUCCESS;
    }
    if (auto *s = dynamic_cast<SolverOp<float> *>(cg->impl.get())) {
        auto *c = dynamic_cast<sfem::ConjugateGradient<float> *>(s->op.get());
        SFEM_AMD_REQUIRE(c && c->flexible, "cg_add_fallback: needs a flexible CG (mode 1)");
        chain(c->on_iteration);
        s->keep.push_back(swp);
        return SFEM_SUCCESS;
    }
    set_last_error("cg_add_fallback: not a CG solver");
    return SFEM_FAILURE;
}

}  // extern "C"
