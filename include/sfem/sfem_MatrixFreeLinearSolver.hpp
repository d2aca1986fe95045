// sfem::MatrixFreeLinearSolver<T> (algebra/sfem_MatrixFreeLinearSolver.hpp:191-215).
#pragma once

#include "sfem_Operator.hpp"

namespace sfem {

template <typename T>
class MatrixFreeLinearSolver : public Operator<T> {
public:
    virtual ~MatrixFreeLinearSolver() = default;
    virtual void set_op(const SharedOperator<T> &op) = 0;
    virtual void set_preconditioner_op(const SharedOperator<T> &op) = 0;
    virtual void set_max_it(int it) = 0;
    virtual void set_n_dofs(std::ptrdiff_t n) = 0;
    virtual void set_initial_guess_zero(bool) {}
    // Extension: with set_initial_guess_zero(true) the solver does not read x (the caller may skip
    // zeroing it).
    virtual bool zero_guess_ignores_x() const { return false; }
    virtual int iterations() const = 0;
};

}  // namespace sfem
