// Internal types shared by the runtime translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "transfer_blas_kernels.hpp"

namespace sfem_amd {

// ----------------------------------------------------------------- device memory helper
template <typename T>
struct DevArray {
    T *p{nullptr};
    size_t n{0};
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        reset();
        if (count == 0) return SFEM_SUCCESS;
        SFEM_AMD_HIP_CHECK(hipMalloc((void **)&p, sizeof(T) * count));
        n = count;
        return SFEM_SUCCESS;
    }
    int upload(const T *h, size_t count) {
        if (alloc(count)) return SFEM_FAILURE;
        // host or device source (unified addressing)
        if (count) SFEM_AMD_HIP_CHECK(hipMemcpy(p, h, sizeof(T) * count, hipMemcpyDefault));
        return SFEM_SUCCESS;
    }
};

inline double *reduction_ws() {
    static double *ws = nullptr;
    if (!ws) {
        if (hipMalloc((void **)&ws, sizeof(double) * reduction_workspace_doubles()) != hipSuccess) ws = nullptr;
    }
    return ws;
}

inline int fill_zero(void *p, size_t bytes, hipStream_t s) {
    return blas_values<float>((ptrdiff_t)(bytes / sizeof(float)), 0.f, (float *)p, s);
}

inline uint16_t to_bf16(float f) {
    // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

// ----------------------------------------------------------------- operator base
struct OpBase {
    virtual ~OpBase() = default;
    virtual int apply(const void *x, void *y, hipStream_t s) = 0;
    // y = A x (overwriting y).  Default: zero y, then accumulate; operators whose kernels own
    // every output entry override it and skip both the fill and the read of y.
    virtual int apply_overwrite(const void *x, void *y, hipStream_t s) {
        if (fill_zero(y, (size_t)rows() * (real_type() == SFEM_AMD_FLOAT64 ? 8 : 4), s)) return SFEM_FAILURE;
        return apply(x, y, s);
    }
    virtual ptrdiff_t rows() const = 0;
    virtual ptrdiff_t cols() const { return rows(); }
    virtual int real_type() const = 0;
    virtual int diag(void *, hipStream_t) {
        set_last_error("diag not available for this operator");
        return SFEM_FAILURE;
    }
    virtual int block_diag(void *, hipStream_t) {
        set_last_error("block diagonal not available for this operator");
        return SFEM_FAILURE;
    }
    // sshex8 operators: the mesh whose skeleton gather writes their shared nodes (Dirichlet rows on
    // it can be set there, see apply_constrained); null otherwise
    virtual const sfem_amd_mesh *gather_mesh(hipStream_t) { return nullptr; }
    virtual int block_size() const { return 1; }
    // y (+)= A x with the rows flagged in cmask (per entry of gather_mesh's skeleton, bit per
    // component) set to x instead; overwrite: y is output only
    virtual int apply_constrained(const void *, void *, const uint8_t *, int, hipStream_t) {
        set_last_error("constrained gather not available for this operator");
        return SFEM_FAILURE;
    }
    // Prolongations: y += P x with constrained fine rows unchanged (the multigrid's coarse
    // correction in one pass), where adds_correction says so.
    virtual bool adds_correction(hipStream_t) { return false; }
    virtual int apply_add(const void *, void *, hipStream_t) {
        set_last_error("adding prolongation not available for this operator");
        return SFEM_FAILURE;
    }
    // A Chebyshev3 step run inside the apply (ChebStep: with t = A x, p and x updated in place, t
    // not stored); cmask as in apply_constrained.  Only where cheb3_fusable says so.
    virtual bool cheb3_fusable(hipStream_t) { return false; }
    virtual int cheb3_step(void *, const void *, const void *, void *, double, double, int, const uint8_t *,
                           hipStream_t) {
        set_last_error("fused Chebyshev step not available for this operator");
        return SFEM_FAILURE;
    }
    // Solvers and smoothers: the output holds zeros on entry to the next applies (work that
    // multiplies the initial guess may be skipped).  Other operators ignore it.
    virtual void set_initial_guess_zero(bool) {}
    // with set_initial_guess_zero(true) the output's contents are not read
    virtual bool zero_guess_ignores_x() const { return false; }
    virtual int iterations() const { return -1; }
    virtual int last_status() const { return SFEM_SUCCESS; }
    virtual const std::vector<double> *history() const { return nullptr; }
};

}  // namespace sfem_amd

namespace sfem_amd {
struct SkeletonStore {
    DevArray<int32_t> node, off, pos, ent;
    ptrdiff_t n{0};
    bool covers{false};  // every mesh node lies in some element (overwriting applies allowed)
};

// The elements of a tet4/hex8 mesh grouped into patches of at most kP elements, each grown
// breadth-first over shared nodes from a seed next to the previous patch (mesh_patches).  A patch
// kernel gathers the patch's nodes once into LDS, accumulates its elements there, adds the nodes that
// belong to this patch alone straight into y, and writes the others (the patch boundary) to
// partials in node-sorted order, which skeleton_gather sums per node in patch order.
struct ElementPatches {
    static constexpr int kP = 512;
    ptrdiff_t npatch{0};
    int maxloc{0};                   // largest node count of a patch
    double nodes_per_element{0};     // patch node entries / elements (1 would be no duplication)
    DevArray<int32_t> count;         // [npatch] elements in each patch (its first slots)
    DevArray<int32_t> perm;          // [npatch * kP] element in each slot (-1 past the patch's end)
    DevArray<uint16_t> lconn;        // [npatch][nv][kP] local node of each element corner
    DevArray<int32_t> node_off;      // [npatch + 1] ranges of the patch node lists
    DevArray<int32_t> nodes;         // mesh node of each patch node
    DevArray<int32_t> bslot;         // its partial position (node-sorted), -1 for a patch's own node
    SkeletonStore boundary;          // the shared nodes: node (ascending), off (partial ranges)
    ptrdiff_t nparts{0};             // partial entries
};
}  // namespace sfem_amd

struct sfem_amd_op {
    std::shared_ptr<sfem_amd::OpBase> impl;
};

struct sfem_amd_mesh {
    int type{0};
    int level{1};
    ptrdiff_t E{0};
    ptrdiff_t nnodes{0};
    int nv{0};                             // nodes per element
    sfem_amd::DevArray<int32_t> idx;       // SoA [nv][E] (tet4/hex8) or element-major [E][nv] (sshex8)
    mutable sfem_amd::DevArray<uint16_t> count;  // incidence counts (lazy)
    mutable sfem_amd::DevArray<uint8_t> own;     // owner mask (lazy)
    const sfem_amd_mesh *parent{nullptr};        // set by sfem_amd_mesh_derefine
    // sshex8: per-node skeleton slot lists of the mesh viewed at level Lc (lazy, mesh_skeleton)
    mutable std::map<int, std::unique_ptr<sfem_amd::SkeletonStore>> skel;
    // sshex8: the macro elements cut into sub-blocks of level Ls (lazy, mesh_subblocks)
    mutable std::map<int, std::unique_ptr<sfem_amd_mesh>> sub;
    // tet4/hex8: element patches (lazy, mesh_patches)
    mutable std::unique_ptr<sfem_amd::ElementPatches> patches;
};

namespace sfem_amd {
// Library options (sfem_amd_set_option): Chebyshev3 steps fused into the sshex8 matrix-core
// elasticity applies (default off: measured slower than the streaming update kernels).
extern bool g_fused_chebyshev;
// sshex8 Laplacian accumulating applies sum skeleton nodes with atomics (option "atomic_skeleton")
extern bool g_atomic_skeleton;
// sshex8 elasticity operators: 0 stencil form for FP64 at levels >= 10, element form otherwise; 1
// element form (matrix-core / node-wise kernels) everywhere; 2 stencil form wherever supported
// (option "elasticity_form")
extern int g_elasticity_form;
// hex8 elasticity applies on element patches (option "hex8_patches", default on) for meshes of at
// least kPatchMinElements elements; off: one element per lane with atomics
extern bool g_hex8_patches;
constexpr ptrdiff_t kPatchMinElements = (ptrdiff_t)1 << 15;
// Host-side state that a captured HIP graph depends on (library options, switch selections, CNN
// weights) changes the state epoch; graph-replaying operators re-capture when it moved.
uint64_t state_epoch();
void bump_state_epoch();
// Zeroes / sets constrained dofs; defined in runtime.cpp.
int constrained_set_value(const OpBase *constrained, int rt, double v, void *x, hipStream_t s);
int mesh_counts(const sfem_amd_mesh *m, hipStream_t s);
bool is_constrained(const OpBase *op);
// Builds (once) the skeleton slot lists of an sshex8 mesh from its connectivity.
int mesh_skeleton(const sfem_amd_mesh *m, int Lc);
// The sshex8 mesh's macro elements as (L/Ls)^3 sub-blocks of level Ls (sshex8_subblock_idx), built
// once per level and shared; null on failure.
const sfem_amd_mesh *mesh_subblocks(const sfem_amd_mesh *m, int Ls, hipStream_t s);
// Builds (once) the element patches of a tet4/hex8 mesh; null on failure or for other meshes.
const ElementPatches *mesh_patches(const sfem_amd_mesh *m);
sfem_amd_op *wrap_op(std::shared_ptr<OpBase> o);
}  // namespace sfem_amd
