// sfem::create_ssgmg (ssmg/sfem_ssgmg.cpp:9-42) with the reference's default hierarchy
// (create_gmg_data, create_gmg_operators, create_gmg_default_smoothers_and_solver:
// algebra/mg/sfem_GeometricMultigrid.cpp:12-168) on the device, over the library's C ABI:
//  * levels: the derefinement levels of the fine sshex8 level L (L, L/2, ... while the quotient stays
//    even, then 1: 16 -> 16 8 4 2 1), level 1 being the HEX8 macro mesh; every level rediscretises
//    the operator with the macro geometry; coarse constrained nodes are the fine ones that exist
//    there (hierarchical numbering: every level is a prefix of the fine numbering);
//  * restriction and prolongation zero the constrained dofs of their output level;
//  * smoothers: 3 sweeps of StationaryIteration with Jacobi, omega 1 (scalar problems), or with the
//    symmetric 3x3 nodal-block Jacobi, omega 1/3 (elasticity);
//  * coarsest level: ConjugateGradient, rtol 1e-6, at most 10000 iterations, preconditioned by the
//    inverse diagonal scaled by 1/block_size.
// The reference builds this from an sfem::Function; the drop-in takes the same content as plain
// arrays (SSGMGProblem).  As the reference's create_ssgmg, the result is a Multigrid with its
// default max_it and atol; set_max_it(1) makes it a one-cycle preconditioner.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "sfem_API.hpp"

namespace sfem {

struct SSGMGProblem {
    const sfem_amd_mesh *mesh{nullptr};  // fine sshex8 mesh; must outlive the returned multigrid
    enum Kind { LAPLACIAN = 0, LINEAR_ELASTICITY = 1 } kind{LAPLACIAN};
    const float *fff{nullptr};                              // Laplacian: macro FFF [E][6] (host or device)
    const float *adjugate{nullptr}, *determinant{nullptr};  // elasticity: macro [E][9], [E]
    double mu{0}, lambda{0};
    const float *coeff{nullptr};                 // optional per-macro-element multiplier [E]
    const int32_t *dirichlet_nodes{nullptr};     // host node ids (fine numbering)
    std::ptrdiff_t n_dirichlet{0};
};

// Levels of the hierarchical numbering, finest first (sshex8_hierarchical_mesh_levels,
// mesh/sfem_hex8_mesh_graph.cpp:964-992; frontend/tests/sfem_SemiStructuredMeshTest.cpp:18-22).
inline std::vector<int> derefinement_levels(int L) {
    std::vector<int> out{L};
    while (out.back() % 2 == 0 && (out.back() / 2) % 2 == 0) out.push_back(out.back() / 2);
    if (out.back() != 1) out.push_back(1);
    return out;
}

namespace detail {
// zero-filled device memory (the diagonal assemblies accumulate into their output); the fill runs on
// the library stream, which the assemblies below also use
struct DeviceBuffer {
    void *p{nullptr};
    explicit DeviceBuffer(size_t bytes) : p(sfem_amd_malloc(bytes ? bytes : 1)) {
        if (!p || sfem_amd_memset(p, 0, bytes, nullptr) != SFEM_SUCCESS) fatal("create_ssgmg: device buffer");
    }
    ~DeviceBuffer() { sfem_amd_free(p); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};
inline void check(int err, const char *what) {
    if (err != SFEM_SUCCESS) fatal(what);
}
}  // namespace detail

inline std::shared_ptr<Multigrid<real_t>> create_ssgmg(const SSGMGProblem &p, const ExecutionSpace es) {
    require_device(es, "create_ssgmg");
    if (!p.mesh || sfem_amd_mesh_element_type(p.mesh) != SFEM_AMD_SSHEX8) fatal("create_ssgmg: sshex8 mesh required");
    const int bs = p.kind == SSGMGProblem::LINEAR_ELASTICITY ? 3 : 1;
    const std::vector<int> levels = derefinement_levels(sfem_amd_mesh_level(p.mesh));
    const int nlevels = (int)levels.size();

    // the coarse meshes live as long as any operator made on them
    struct Meshes {
        std::vector<sfem_amd_mesh *> owned;
        ~Meshes() {
            for (auto *m : owned) sfem_amd_mesh_destroy(m);
        }
    };
    auto meshes = std::make_shared<Meshes>();
    std::vector<const sfem_amd_mesh *> mesh{p.mesh};
    for (int l = 1; l < nlevels; l++) {
        sfem_amd_mesh *m = nullptr;
        detail::check(sfem_amd_mesh_derefine(mesh.back(), levels[l], &m), "create_ssgmg: mesh_derefine");
        meshes->owned.push_back(m);
        mesh.push_back(m);
    }

    std::vector<sfem_amd_op *> op(nlevels), A(nlevels);
    for (int l = 0; l < nlevels; l++) {
        if (bs == 1)
            detail::check(sfem_amd_laplacian_create(mesh[l], p.fff, p.coeff, SFEM_AMD_FLOAT64, &op[l]),
                          "create_ssgmg: laplacian_create");
        else
            detail::check(sfem_amd_linear_elasticity_create(mesh[l], p.adjugate, p.determinant, p.mu, p.lambda,
                                                            p.coeff, SFEM_AMD_FLOAT64, &op[l]),
                          "create_ssgmg: linear_elasticity_create");
        const std::ptrdiff_t nn = sfem_amd_mesh_nnodes(mesh[l]);
        std::vector<int32_t> bc;
        for (std::ptrdiff_t i = 0; i < p.n_dirichlet; i++)
            if (p.dirichlet_nodes[i] < nn) bc.push_back(p.dirichlet_nodes[i]);
        detail::check(sfem_amd_constrained_op_create(op[l], (std::ptrdiff_t)bc.size(), bc.data(), bs, 0, nullptr, &A[l]),
                      "create_ssgmg: constrained_op_create");
    }

    // one owning view per constrained operator handle
    std::vector<SharedOperator<real_t>> Aop(nlevels);
    for (int l = 0; l < nlevels; l++) Aop[l] = make_device_op<real_t>(A[l], nullptr, meshes);

    auto mg = create_mg<real_t>(es);
    for (int l = 0; l < nlevels; l++) {
        const std::ptrdiff_t n = sfem_amd_op_rows(A[l]);
        sfem_amd_op *smoother = nullptr;
        SharedOperator<real_t> sm;
        if (l < nlevels - 1) {
            if (bs == 1) {
                detail::DeviceBuffer d(sizeof(real_t) * n);
                detail::check(sfem_amd_op_diag(A[l], d.p, nullptr), "create_ssgmg: diag");
                detail::check(sfem_amd_jacobi_create(A[l], d.p, 1.0, 3, &smoother), "create_ssgmg: jacobi_create");
                detail::check(sfem_amd_device_synchronize(), "create_ssgmg: synchronize");
            } else {
                detail::DeviceBuffer d6(sizeof(real_t) * (n / 3) * 6), mask(n);
                detail::check(sfem_amd_op_block_diag_sym(op[l], d6.p, nullptr), "create_ssgmg: block_diag_sym");
                // constrained dofs: their block rows become unit rows
                std::vector<uint8_t> hm((size_t)n, 0);
                {
                    std::vector<real_t> ones((size_t)n, 1.0), probe((size_t)n, 0.0);
                    detail::DeviceBuffer x(sizeof(real_t) * n);
                    detail::check(sfem_amd_memcpy_h2d(x.p, ones.data(), sizeof(real_t) * n), "create_ssgmg: h2d");
                    detail::check(sfem_amd_constrained_op_set_value(A[l], 0.0, x.p, nullptr), "create_ssgmg: set_value");
                    detail::check(sfem_amd_device_synchronize(), "create_ssgmg: synchronize");
                    detail::check(sfem_amd_memcpy_d2h(probe.data(), x.p, sizeof(real_t) * n), "create_ssgmg: d2h");
                    for (std::ptrdiff_t i = 0; i < n; i++) hm[(size_t)i] = probe[(size_t)i] == 0.0 ? 1 : 0;
                }
                detail::check(sfem_amd_memcpy_h2d(mask.p, hm.data(), (size_t)n), "create_ssgmg: h2d mask");
                detail::check(sfem_amd_block_jacobi_create(A[l], d6.p, (const uint8_t *)mask.p, 1.0 / 3, 3, &smoother),
                              "create_ssgmg: block_jacobi_create");
                detail::check(sfem_amd_device_synchronize(), "create_ssgmg: synchronize");
            }
            sm = make_device_op<real_t>(smoother, nullptr, meshes);
        } else {
            // coarsest: PCG, rtol 1e-6, inverse diagonal scaled by 1 / block_size
            detail::DeviceBuffer d(sizeof(real_t) * n);
            detail::check(sfem_amd_op_diag(A[l], d.p, nullptr), "create_ssgmg: coarse diag");
            auto cg = create_cg<real_t>(Aop[l], es);
            cg->set_preconditioner_op(create_shiftable_jacobi<real_t>((const real_t *)d.p, n, real_t(1) / bs, es));
            cg->set_max_it(10000);
            cg->set_rtol(1e-6);
            detail::check(sfem_amd_device_synchronize(), "create_ssgmg: synchronize");
            sm = cg;
        }
        SharedOperator<real_t> prolongation, restriction;
        if (l > 0) {
            sfem_amd_op *P = nullptr;
            detail::check(sfem_amd_prolongation_create(mesh[l], mesh[l - 1], bs, SFEM_AMD_FLOAT64, A[l - 1], &P),
                          "create_ssgmg: prolongation_create");
            prolongation = make_device_op<real_t>(P, nullptr, meshes);
        }
        if (l < nlevels - 1) {
            sfem_amd_op *R = nullptr;
            detail::check(sfem_amd_restriction_create(mesh[l], mesh[l + 1], bs, SFEM_AMD_FLOAT64, A[l + 1], &R),
                          "create_ssgmg: restriction_create");
            restriction = make_device_op<real_t>(R, nullptr, meshes);
        }
        mg->add_level(Aop[l], sm, prolongation, restriction);
    }
    for (auto *o : op) sfem_amd_op_destroy(o);  // the constrained operators keep them
    return mg;
}

}  // namespace sfem
