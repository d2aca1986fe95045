// sfem::Chebyshev3<T> (algebra/sfem_Chebyshev3.hpp:57-303): three-term Chebyshev
// smoother over [scale_eig_min*lmax, lmax], lmax = scale_eig_max * eig_max.
// Extensions (not in the reference): with `inv_diag` set before init(), the recurrence and the
// eigenvalue estimate run on D^-1 A (Jacobi-scaled Chebyshev), for coefficients that vary between
// elements by orders of magnitude; with `lanczos_steps` > 0, init() takes eig_max from at most
// that many Lanczos steps (relative tolerance eigen_solver_tol) instead of the power method.
#pragma once

#include "sfem_Buffer.hpp"
#include "sfem_MatrixFreeLinearSolver.hpp"
#include "sfem_Lanczos.hpp"
#include "sfem_PowerMethod.hpp"

namespace sfem {

// Extension: an operator that can run a Chebyshev3 step inside its apply.  With t = A x:
// p = D^-1 (t - rhs) (first step) or beta p + D^-1 (t - rhs), then x -= alpha p, without storing t
// (D^-1 = inv_diag, or the identity when null).  The smoother uses it when cheb3_fusable().
template <typename T>
class FusedChebyshevStep {
public:
    virtual ~FusedChebyshevStep() = default;
    virtual bool cheb3_fusable() = 0;
    virtual int cheb3_step(T *x, const T *rhs, const T *inv_diag, T *p, T alpha, T beta, int first) = 0;
};

template <typename T>
class Chebyshev3 final : public MatrixFreeLinearSolver<T> {
public:
    SharedOperator<T> op;
    std::shared_ptr<BLAS<T>> blas;
    std::shared_ptr<PowerMethod<T>> power_method;
    SharedBuffer<T> p_, temp_;
    SharedBuffer<T> inv_diag;  // optional D^-1 (device)

    T eigen_solver_tol{1e-6};
    int eigen_solver_max_it{1000};
    int lanczos_steps{0};  // > 0: eig_max by Lanczos (extension)
    int eigen_iterations{0};
    int max_it{3};
    int iterations_{0};
    T eig_max{0};
    T scale_eig_max{1};
    T scale_eig_min{0.06};
    std::ptrdiff_t n_dofs{-1};
    bool is_initial_guess_zero{false};
    bool apply_overwrites_output{false};
    void *stream{nullptr};

    void default_init(void *s = nullptr) {
        stream = s;
        blas = make_device_blas<T>(s);
        power_method = std::make_shared<PowerMethod<T>>();
        power_method->blas = blas;
    }
    int iterations() const override { return iterations_; }
    void set_initial_guess_zero(bool v) override { is_initial_guess_zero = v; }
    bool zero_guess_ignores_x() const override { return true; }  // the first step writes x
    void set_op(const SharedOperator<T> &o) override {
        op = o;
        n_dofs = o->rows();
    }
    void set_preconditioner_op(const SharedOperator<T> &) override { fatal("Chebyshev3 is unpreconditioned"); }
    void set_max_it(int it) override { max_it = it; }
    void set_n_dofs(std::ptrdiff_t n) override { n_dofs = n; }
    std::ptrdiff_t rows() const override { return n_dofs; }
    std::ptrdiff_t cols() const override { return n_dofs; }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }

    void init_with_ones() {
        auto ones = Buffer<T>::own(n_dofs, blas->allocate(n_dofs), [b = blas](void *q) { b->destroy(q); });
        blas->values(n_dofs, 1, ones->data());
        init(ones->data());
    }
    // algebra/sfem_Chebyshev3.hpp:165-175
    void init(const T *guess) {
        T *ev = blas->allocate(n_dofs);
        T *work = blas->allocate(n_dofs);
        blas->copy(n_dofs, guess, ev);
        auto apply_fn = [this](const T *const x, T *const y) {
            op->apply(x, y);
            if (inv_diag) blas->xypaz(n_dofs, inv_diag->data(), y, T(0), y);
        };
        if (lanczos_steps > 0) {
            Lanczos<T> lanczos;
            lanczos.blas = blas;
            eig_max = lanczos.max_eigen_value([this](const T *const x, T *const y) { op->apply(x, y); },
                                              inv_diag ? inv_diag->data() : nullptr, lanczos_steps, eigen_solver_tol,
                                              n_dofs, guess);
            eigen_iterations = lanczos.last_iterations;
        } else {
            eig_max = power_method->max_eigen_value(apply_fn, eigen_solver_max_it, eigen_solver_tol, n_dofs, ev, work);
            eigen_iterations = power_method->last_iterations;
        }
        auto b = blas;
        p_ = Buffer<T>::own(n_dofs, work, [b](void *q) { b->destroy(q); });
        temp_ = Buffer<T>::own(n_dofs, ev, [b](void *q) { b->destroy(q); });
    }

    int apply(const T *const rhs, T *const x) override {
        if (!op || !p_) return SFEM_FAILURE;
        if (max_it <= 0) {
            iterations_ = 0;
            return SFEM_SUCCESS;
        }
        const std::ptrdiff_t n = n_dofs;
        T *p = p_->data(), *temp = temp_->data();
        const T lmax = eig_max * scale_eig_max;
        const T lmin = scale_eig_min * lmax;
        const T avg = (lmin + lmax) / 2, diff = (lmin - lmax) / 2;
        T alpha = T(1) / avg, beta = 0, dea = 0;
        // operators that run the step inside their apply (same words, no product vector)
        auto *fz = dynamic_cast<FusedChebyshevStep<T> *>(op.get());
        const bool fused = fz && apply_overwrites_output && fz->cheb3_fusable();
        const T *dinv = inv_diag ? inv_diag->data() : nullptr;
        // iteration 0: p = A x - rhs (or -rhs when x == 0), scaled by D^-1 when given; x -= alpha p
        // (one fused pass after the product)
        if (!is_initial_guess_zero && fused) {
            if (fz->cheb3_step(x, rhs, dinv, p, alpha, T(0), 1)) return SFEM_FAILURE;
        } else {
            if (!is_initial_guess_zero) {
                if (!apply_overwrites_output) blas->zeros(n, p);
                op->apply(x, p);
            }
            sfem_amd_cheb3_start(real_type_of<T>::value, n, (double)alpha, rhs, dinv, is_initial_guess_zero ? 0 : 1,
                                 p, x, stream);
        }
        iterations_ = 1;
        for (; iterations_ < max_it; iterations_++) {
            dea = diff * alpha;
            beta = (iterations_ == 1 ? T(0.5) : T(0.25)) * dea * dea;
            alpha = T(1) / (avg - beta / alpha);
            if (fused) {
                if (fz->cheb3_step(x, rhs, dinv, p, alpha, beta, 0)) return SFEM_FAILURE;
                continue;
            }
            if (!apply_overwrites_output) blas->zeros(n, temp);
            op->apply(x, temp);
            // fused: p = beta p - rhs + A x (times D^-1 when scaled); x -= alpha p
            if (inv_diag)
                sfem_amd_cheb3_update_diag(real_type_of<T>::value, n, alpha, beta, rhs, temp, inv_diag->data(), p, x,
                                           stream);
            else
                sfem_amd_cheb3_update(real_type_of<T>::value, n, alpha, beta, rhs, temp, p, x, stream);
        }
        return SFEM_SUCCESS;
    }
};

}  // namespace sfem
