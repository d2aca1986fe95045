// Factories of the reference's frontend/sfem_API.hpp for the device execution space:
// create_cg (:143), create_mg (:287), create_cheb3 (:323), plus create_stationary and
// create_shiftable_jacobi used by the default multigrid smoothers
// (algebra/mg/sfem_GeometricMultigrid.cpp:89-168).  Objects compute on the GPU through the
// library's C ABI; EXECUTION_SPACE_HOST has no implementation here and is a fatal error, as an
// unknown execution space is in the reference.
#pragma once

#include "sfem_Chebyshev3.hpp"
#include "sfem_Multigrid.hpp"
#include "sfem_Stationary.hpp"
#include "sfem_cg.hpp"
#include "sfem_hip_ops.hpp"

namespace sfem {

inline void require_device(ExecutionSpace es, const char *what) {
    if (es != EXECUTION_SPACE_DEVICE) {
        std::fprintf(stderr, "[sfem_amd] %s: only EXECUTION_SPACE_DEVICE is available\n", what);
        std::abort();
    }
}

template <typename T>
std::shared_ptr<ConjugateGradient<T>> create_cg(const SharedOperator<T> &op, const ExecutionSpace es) {
    require_device(es, "create_cg");
    auto cg = std::make_shared<ConjugateGradient<T>>();
    cg->default_init();
    cg->set_n_dofs(op->rows());
    cg->set_op(op);
    return cg;
}

template <typename T>
std::shared_ptr<Chebyshev3<T>> create_cheb3(const SharedOperator<T> &op, const ExecutionSpace es) {
    require_device(es, "create_cheb3");
    auto cheb = std::make_shared<Chebyshev3<T>>();
    cheb->default_init();
    cheb->set_op(op);
    return cheb;
}

template <typename T>
std::shared_ptr<Multigrid<T>> create_mg(const ExecutionSpace es) {
    require_device(es, "create_mg");
    auto mg = std::make_shared<Multigrid<T>>();
    mg->default_init();
    return mg;
}

// The reference's create_shiftable_jacobi(diag, es) with relaxation_parameter omega:
// x += omega D^-1 r (diag: device vector of n entries, copied).
template <typename T>
SharedOperator<T> create_shiftable_jacobi(const T *diag, std::ptrdiff_t n, T omega, const ExecutionSpace es) {
    require_device(es, "create_shiftable_jacobi");
    return std::make_shared<Jacobi<T>>(diag, n, omega, make_device_blas<T>());
}

template <typename T>
std::shared_ptr<StationaryIteration<T>> create_stationary(const SharedOperator<T> &op,
                                                          const SharedOperator<T> &preconditioner,
                                                          const ExecutionSpace es) {
    require_device(es, "create_stationary");
    auto s = std::make_shared<StationaryIteration<T>>();
    s->default_init();
    s->set_op(op);
    s->set_preconditioner_op(preconditioner);
    return s;
}

}  // namespace sfem
