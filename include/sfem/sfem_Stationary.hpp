// sfem::StationaryIteration<T> (algebra/sfem_Stationary.hpp:19-122) and the scaled Jacobi
// preconditioner it is used with (algebra/sfem_ShiftableJacobi.hpp): x += omega D^-1 (b - A x).
#pragma once

#include "sfem_Buffer.hpp"
#include "sfem_MatrixFreeLinearSolver.hpp"
#include "sfem_tpl_blas.hpp"

namespace sfem {

// x += omega * inv_diag * r (applied as an operator r -> x); omega fixed at construction
template <typename T>
class Jacobi final : public Operator<T> {
public:
    SharedBuffer<T> inv_diag;  // omega / diag
    std::shared_ptr<BLAS<T>> blas;
    Jacobi(const T *diag, std::ptrdiff_t n, T omega, std::shared_ptr<BLAS<T>> b) : blas(std::move(b)) {
        inv_diag = Buffer<T>::own(n, blas->allocate(n), [bb = blas](void *q) { bb->destroy(q); });
        blas->copy(n, diag, inv_diag->data());
        blas->reciprocal(n, omega, inv_diag->data());
    }
    int apply(const T *const r, T *const x) override {
        blas->xypaz(inv_diag->size(), inv_diag->data(), r, T(1), x);
        return SFEM_SUCCESS;
    }
    std::ptrdiff_t rows() const override { return inv_diag->size(); }
    std::ptrdiff_t cols() const override { return inv_diag->size(); }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }
};

template <typename T>
class StationaryIteration final : public MatrixFreeLinearSolver<T> {
public:
    SharedOperator<T> op, preconditioner;
    std::shared_ptr<BLAS<T>> blas;
    SharedBuffer<T> workspace;
    std::ptrdiff_t n_dofs{-1};
    int max_it{3};
    int iterations_{0};

    void default_init(void *stream = nullptr) { blas = make_device_blas<T>(stream); }
    int iterations() const override { return iterations_; }
    void set_op(const SharedOperator<T> &o) override {
        op = o;
        n_dofs = o->rows();
    }
    void set_preconditioner_op(const SharedOperator<T> &p) override { preconditioner = p; }
    void set_max_it(int it) override { max_it = it; }
    void set_n_dofs(std::ptrdiff_t n) override { n_dofs = n; }
    std::ptrdiff_t rows() const override { return n_dofs; }
    std::ptrdiff_t cols() const override { return n_dofs; }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }

    int apply(const T *const b, T *const x) override {
        if (!workspace || workspace->size() != n_dofs)
            workspace = Buffer<T>::own(n_dofs, blas->allocate(n_dofs), [bb = blas](void *q) { bb->destroy(q); });
        T *r = workspace->data();
        for (iterations_ = 0; iterations_ < max_it; iterations_++) {
            blas->zeros(n_dofs, r);
            op->apply(x, r);
            blas->axpby(n_dofs, T(1), b, T(-1), r);
            preconditioner->apply(r, x);
        }
        return SFEM_SUCCESS;
    }
};

}  // namespace sfem
