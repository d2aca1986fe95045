// A C++ caller of the drop-in: the reference's device solve path (create_ssgmg + ConjugateGradient on
// device pointers) written against include/sfem/ and the C ABI only.  Built by
// tests/test_gpu_cpp_drop_in.py with g++ and linked to libsfem_amd.so.
//
//   drop_in_example laplacian|elasticity N L OUT
//
// N^3 hex8 macro elements of the unit cube refined to sshex8 level L; Laplacian: zero Dirichlet on
// the whole boundary, f = 1; elasticity: clamped at z = 0, body force (0, 0, -1), E = 1, nu = 0.3.
// CG with one V-cycle of create_ssgmg's default hierarchy as preconditioner, rtol 1e-12.  Writes
// the solution (float64, interleaved) to OUT and prints "iterations <k> status <s>".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sfem/sfem_ssgmg.hpp"

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s laplacian|elasticity N L OUT\n", argv[0]);
        return 2;
    }
    const bool elasticity = std::strcmp(argv[1], "elasticity") == 0;
    const int n = std::atoi(argv[2]), L = std::atoi(argv[3]);
    const int bs = elasticity ? 3 : 1;

    // macro mesh: the box_mesh layout (python/sfem/mesh/box_mesh.py:479-700), host arrays
    const std::ptrdiff_t nn = sfem_amd_box_nnodes(n + 1, n + 1, n + 1);
    const std::ptrdiff_t ne = sfem_amd_box_nelements(n + 1, n + 1, n + 1, SFEM_AMD_HEX8);
    std::vector<int32_t> mel(8 * ne);
    std::vector<float> mpts(3 * nn);
    if (sfem_amd_box_mesh(n + 1, n + 1, n + 1, 1.0, 1.0, 1.0, SFEM_AMD_HEX8, mel.data(), mpts.data())) return 1;
    // sshex8 refinement with the hierarchical numbering (the reference's SoA layout [nxe][E])
    const ptrdiff_t nnodes = sfem_amd_sshex8_refine(L, ne, nn, mel.data(), mpts.data(), nullptr, nullptr);
    const int nxe = (L + 1) * (L + 1) * (L + 1);
    std::vector<int32_t> el((size_t)nxe * ne);
    std::vector<float> pts(3 * (size_t)nnodes);
    sfem_amd_sshex8_refine(L, ne, nn, mel.data(), mpts.data(), el.data(), pts.data());
    sfem_amd_mesh *mesh = nullptr;
    if (sfem_amd_mesh_create(SFEM_AMD_SSHEX8, L, ne, ne, el.data(), nnodes, &mesh)) return 1;

    sfem::SSGMGProblem prob;
    prob.mesh = mesh;
    std::vector<float> fff, adj, det;
    std::vector<int32_t> bc;
    for (ptrdiff_t i = 0; i < nnodes; i++) {
        const float x = pts[i], y = pts[nnodes + i], z = pts[2 * nnodes + i];
        const bool on = elasticity ? z < 1e-6f
                                   : (x < 1e-6f || y < 1e-6f || z < 1e-6f || x > 1 - 1e-6f || y > 1 - 1e-6f ||
                                      z > 1 - 1e-6f);
        if (on) bc.push_back((int32_t)i);
    }
    prob.dirichlet_nodes = bc.data();
    prob.n_dirichlet = (std::ptrdiff_t)bc.size();
    if (elasticity) {
        adj.resize(9 * ne);
        det.resize(ne);
        sfem_amd_adjugate_det(SFEM_AMD_HEX8, 1, ne, ne, mel.data(), nn, mpts.data(), adj.data(), det.data());
        prob.kind = sfem::SSGMGProblem::LINEAR_ELASTICITY;
        prob.adjugate = adj.data();
        prob.determinant = det.data();
        const double nu = 0.3;
        prob.mu = 1.0 / (2 * (1 + nu));
        prob.lambda = nu / ((1 + nu) * (1 - 2 * nu));
    } else {
        fff.resize(6 * ne);
        sfem_amd_fff(SFEM_AMD_HEX8, 1, ne, ne, mel.data(), nn, mpts.data(), fff.data());
        prob.fff = fff.data();
    }

    auto mg = sfem::create_ssgmg(prob, sfem::EXECUTION_SPACE_DEVICE);
    mg->set_max_it(1);  // one V-cycle per preconditioner application
    mg->set_atol(0);
    auto A = mg->operators().front();

    const std::ptrdiff_t ndofs = nnodes * bs;
    std::vector<double> rhs((size_t)ndofs, 0.0);
    for (ptrdiff_t i = 0; i < nnodes; i++) {
        if (elasticity)
            rhs[(size_t)i * 3 + 2] = -1;
        else
            rhs[(size_t)i] = 1;
    }
    double *d_rhs = (double *)sfem_amd_malloc(sizeof(double) * ndofs);
    double *d_x = (double *)sfem_amd_malloc(sizeof(double) * ndofs);
    sfem_amd_memcpy_h2d(d_rhs, rhs.data(), sizeof(double) * ndofs);
    // the Dirichlet rows of the right-hand side zeroed on the device by an in-place operator
    auto zero_bc = sfem::make_constraint_value_op<sfem::real_t>(A, 0.0);
    if (zero_bc->size() != ndofs || zero_bc->apply(d_rhs) != SFEM_SUCCESS) return 1;
    sfem_amd_memset(d_x, 0, sizeof(double) * ndofs, nullptr);

    auto cg = sfem::create_cg<sfem::real_t>(A, sfem::EXECUTION_SPACE_DEVICE);
    cg->set_preconditioner_op(mg);
    cg->set_rtol(1e-12);
    cg->set_max_it(500);
    const int status = cg->apply(ndofs, d_rhs, d_x);  // apply(n, b, x), as the reference's drivers call it

    std::vector<double> x((size_t)ndofs);
    sfem_amd_memcpy_d2h(x.data(), d_x, sizeof(double) * ndofs);
    FILE *f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(x.data(), sizeof(double), x.size(), f) != x.size()) return 1;
    std::fclose(f);
    std::printf("iterations %d status %d\n", cg->iterations(), status);
    sfem_amd_free(d_rhs);
    sfem_amd_free(d_x);
    cg.reset();
    mg.reset();
    A.reset();
    sfem_amd_mesh_destroy(mesh);
    return status;
}
