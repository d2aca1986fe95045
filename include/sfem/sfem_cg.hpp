// sfem::ConjugateGradient<T> (algebra/sfem_cg.hpp:80-446).
//  * plain CG (aux_apply_basic :301-363) and left-preconditioned CG with a fixed
//    preconditioner (aux_apply_precond :365-445), stopping rules as the reference;
//  * flexible PCG (an extension; the reference has none): beta = r_new.(z_new - z_old) /
//    r_old.z_old (Polak-Ribiere form), for preconditioners that are not a fixed SPD map
//    (FP32 or learned V-cycles).  It stops on the true relative residual ||b - A x||/||b||,
//    and recomputes r = b - A x in FP64 every `residual_replacement` iterations and at
//    the end to guard against drift.
#pragma once

#include <cmath>
#include <functional>
#include <vector>

#include "sfem_MatrixFreeLinearSolver.hpp"
#include "sfem_tpl_blas.hpp"

namespace sfem {

template <typename T>
class ConjugateGradient final : public MatrixFreeLinearSolver<T> {
public:
    SharedOperator<T> apply_op;
    SharedOperator<T> preconditioner_op;
    std::shared_ptr<BLAS<T>> blas;
    void *stream{nullptr};

    T rtol{1e-10};
    T atol{1e-16};
    int max_it{10000};
    int check_each{1};
    std::ptrdiff_t n_dofs{-1};
    int iterations_{0};
    bool verbose{false};
    bool apply_overwrites_output{false};
    bool preconditioner_overwrites_output{false};  // (extension) see precondition()
    bool flexible{false};
    // fixed-preconditioner PCG that stops on ||r||/||b|| (recursive residual) instead of the
    // reference's sqrt(r.z)/sqrt(r0.z0)
    bool stop_on_residual_norm{false};
    int residual_replacement{50};
    int last_status{SFEM_FAILURE};
    std::vector<double> history;  // relative residual per iteration
    // flexible PCG only: called after every iteration with (iteration, previous residual norm,
    // new residual norm); used to switch a preconditioner component that stops converging
    std::function<void(int, double, double)> on_iteration;

    ~ConjugateGradient() { release(); }

    void default_init(void *s = nullptr) {
        stream = s;
        blas = make_device_blas<T>(s);
    }
    int iterations() const override { return iterations_; }
    void set_atol(T v) { atol = v; }
    void set_rtol(T v) { rtol = v; }
    void set_verbose(bool v) { verbose = v; }
    // verbose output every check_each iterations (and the first), as algebra/sfem_cg.hpp:187
    bool report(int it) const { return it == 1 || check_each <= 1 || it % check_each == 0; }
    void set_op(const SharedOperator<T> &op) override {
        apply_op = op;
        n_dofs = op->rows();
    }
    void set_preconditioner_op(const SharedOperator<T> &op) override { preconditioner_op = op; }
    void set_max_it(int it) override { max_it = it; }
    void set_n_dofs(std::ptrdiff_t n) override { n_dofs = n; }
    std::ptrdiff_t rows() const override { return n_dofs; }
    std::ptrdiff_t cols() const override { return n_dofs; }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }

    // apply(n, b, x) (algebra/sfem_cg.hpp:193): the same solve on the first n entries
    int apply(const std::ptrdiff_t n, const T *const b, T *const x) {
        const std::ptrdiff_t keep = n_dofs;
        n_dofs = n;
        const int ret = apply(b, x);
        n_dofs = keep;
        return ret;
    }
    int apply(const T *const b, T *const x) override {
        if (n_dofs < 0 || !apply_op || !blas) return SFEM_FAILURE;
        history.clear();
        int ret;
        if (!preconditioner_op)
            ret = basic(n_dofs, b, x);
        else if (flexible)
            ret = flexible_pcg(n_dofs, b, x);
        else
            ret = pcg(n_dofs, b, x);
        last_status = ret;
        return ret;
    }

private:
    T *r_{nullptr}, *z_{nullptr}, *p_{nullptr}, *Ap_{nullptr};
    std::ptrdiff_t wn_{-1};

    void release() {
        for (T **q : {&r_, &z_, &p_, &Ap_}) {
            if (*q && blas) blas->destroy(*q);
            *q = nullptr;
        }
        wn_ = -1;
    }
    void ensure_workspace(std::ptrdiff_t n) {
        if (wn_ == n) return;
        release();
        r_ = blas->allocate(n);
        z_ = blas->allocate(n);
        p_ = blas->allocate(n);
        Ap_ = blas->allocate(n);
        wn_ = n;
    }
    void zero_for_apply(std::ptrdiff_t n, T *y) {
        if (!apply_overwrites_output) blas->zeros(n, y);
    }
    T update_x_r(std::ptrdiff_t n, T alpha, const T *p, const T *Ap, T *x, T *r, bool with_rtr) {
        return (T)sfem_amd_cg_update_x_r(real_type_of<T>::value, n, alpha, p, Ap, x, r, with_rtr ? 1 : 0, stream);
    }
    // returns (r, r) of the updated r and its (r, z) in rz
    T update_x_r_dot(std::ptrdiff_t n, T alpha, const T *p, const T *Ap, T *x, T *r, const T *z, T &rz) {
        double d = 0;
        const T rtr = (T)sfem_amd_cg_update_x_r_dot(real_type_of<T>::value, n, alpha, p, Ap, x, r, z, &d, stream);
        rz = (T)d;
        return rtr;
    }
    // z = P r from z = 0 whatever the operator does; a solver or multigrid preconditioner is told
    // that its initial guess is zero for this apply (it may skip products with it).  With
    // preconditioner_overwrites_output the operator's apply writes all of z itself (from a zero
    // start), and z is not cleared first.
    void precondition(std::ptrdiff_t n, const T *r, T *z) {
        if (!preconditioner_overwrites_output) blas->zeros(n, z);
        auto *mf = dynamic_cast<MatrixFreeLinearSolver<T> *>(preconditioner_op.get());
        if (mf) mf->set_initial_guess_zero(true);
        preconditioner_op->apply(r, z);
        if (mf) mf->set_initial_guess_zero(false);
    }
    void residual(std::ptrdiff_t n, const T *b, const T *x, T *r) {
        zero_for_apply(n, r);
        apply_op->apply(x, r);
        blas->axpby(n, T(1), b, T(-1), r);
    }

    // algebra/sfem_cg.hpp:301-363
    int basic(std::ptrdiff_t n, const T *b, T *x) {
        ensure_workspace(n);
        T *r = r_, *p = p_, *Ap = Ap_;
        residual(n, b, x, r);
        const T rtr0 = blas->dot(n, r, r);
        const T r_norm0 = std::sqrt(rtr0);
        iterations_ = 0;
        if (rtr0 == 0) return SFEM_SUCCESS;
        blas->copy(n, r, p);
        T rtr = rtr0;
        for (iterations_ = 0; iterations_ < max_it; iterations_++) {
            zero_for_apply(n, Ap);
            apply_op->apply(p, Ap);
            const T ptAp = blas->dot(n, p, Ap);
            const T alpha = rtr / ptAp;
            const T rtr_new = update_x_r(n, alpha, p, Ap, x, r, true);
            const T beta = rtr_new / rtr;
            rtr = rtr_new;
            blas->axpby(n, T(1), r, beta, p);
            const T r_norm = std::sqrt(rtr_new);
            history.push_back((double)(r_norm / r_norm0));
            if (verbose && report(iterations_ + 1)) std::printf("%d: residual abs: %g, rel: %g\n", iterations_ + 1, (double)r_norm, (double)(r_norm / r_norm0));
            // the reference breaks without incrementing its counter (:351-355)
            if (r_norm < atol || rtr_new == 0 || r_norm / r_norm0 < rtol) return SFEM_SUCCESS;
        }
        return SFEM_FAILURE;
    }

    // algebra/sfem_cg.hpp:365-445 (stops on the preconditioned residual sqrt(r.z))
    int pcg(std::ptrdiff_t n, const T *b, T *x) {
        ensure_workspace(n);
        T *r = r_, *z = z_, *p = p_, *Ap = Ap_;
        residual(n, b, x, r);
        const T rtr0 = blas->dot(n, r, r);
        iterations_ = 0;
        if (rtr0 == 0) return SFEM_SUCCESS;
        precondition(n, r, z);
        blas->copy(n, z, p);
        zero_for_apply(n, Ap);
        apply_op->apply(p, Ap);
        const T rtz0 = blas->dot(n, r, z);
        T rtz = rtz0;
        if (rtz == 0) return SFEM_SUCCESS;
        {
            const T ptAp = blas->dot(n, p, Ap);
            update_x_r(n, rtz / ptAp, p, Ap, x, r, false);
        }
        const T bnorm = stop_on_residual_norm ? blas->norm2(n, b) : T(1);
        for (iterations_ = 0; iterations_ < max_it; iterations_++) {
            precondition(n, r, z);
            const T rtz_new = blas->dot(n, r, z);
            // exact convergence (e.g. a one-unknown coarse problem): the next step would be 0/0
            if (rtz_new == 0) return SFEM_SUCCESS;
            const T beta = rtz_new / rtz;
            rtz = rtz_new;
            blas->axpby(n, T(1), z, beta, p);
            zero_for_apply(n, Ap);
            apply_op->apply(p, Ap);
            const T ptAp = blas->dot(n, p, Ap);
            const T alpha = rtz / ptAp;
            const T rtr = update_x_r(n, alpha, p, Ap, x, r, stop_on_residual_norm);
            if (stop_on_residual_norm) {
                const T rel = std::sqrt(rtr) / (bnorm > 0 ? bnorm : T(1));
                history.push_back((double)rel);
                if (rel < rtol || std::sqrt(rtr) < atol) return SFEM_SUCCESS;
                continue;
            }
            const T anorm = std::sqrt(std::abs(rtz));
            const T rnorm = anorm / std::sqrt(rtz0);
            history.push_back((double)rnorm);
            if (verbose && report(iterations_ + 1)) std::printf("%d: sqrt(r.z) abs: %g, rel: %g\n", iterations_ + 1, (double)anorm, (double)rnorm);
            if (anorm < atol || rnorm < rtol) return SFEM_SUCCESS;
        }
        return SFEM_FAILURE;
    }

    // Flexible PCG with true-residual stopping (extension)
    int flexible_pcg(std::ptrdiff_t n, const T *b, T *x) {
        ensure_workspace(n);
        T *r = r_, *z = z_, *p = p_, *Ap = Ap_;
        const T bnorm = blas->norm2(n, b);
        residual(n, b, x, r);
        T rnorm = blas->norm2(n, r);
        iterations_ = 0;
        const T ref = bnorm > 0 ? bnorm : T(1);
        if (rnorm == 0 || rnorm / ref < rtol || rnorm < atol) return SFEM_SUCCESS;
        precondition(n, r, z);
        blas->copy(n, z, p);
        T rtz = blas->dot(n, r, z);
        for (iterations_ = 0; iterations_ < max_it; iterations_++) {
            zero_for_apply(n, Ap);
            apply_op->apply(p, Ap);
            const T ptAp = blas->dot(n, p, Ap);
            const T alpha = rtz / ptAp;
            const bool replace = residual_replacement > 0 && (iterations_ + 1) % residual_replacement == 0;
            // r_new . z_old for beta, taken while z still holds z_old
            T rtzold = 0;
            T rtr = replace ? update_x_r(n, alpha, p, Ap, x, r, false) : update_x_r_dot(n, alpha, p, Ap, x, r, z, rtzold);
            if (replace) {
                residual(n, b, x, r);
                rtr = blas->dot(n, r, r);
                rtzold = blas->dot(n, r, z);
            }
            const T rprev = rnorm;
            rnorm = std::sqrt(rtr);
            history.push_back((double)(rnorm / ref));
            if (verbose && report(iterations_ + 1)) std::printf("%d: residual rel: %g\n", iterations_ + 1, (double)(rnorm / ref));
            if (on_iteration) on_iteration(iterations_, (double)rprev, (double)rnorm);
            if (rnorm / ref < rtol || rnorm < atol) {
                // confirm with the true residual (recomputed in the operator's precision)
                residual(n, b, x, r);
                const T true_norm = blas->norm2(n, r);
                if (true_norm / ref < rtol || true_norm < atol) {
                    iterations_++;
                    if (!history.empty()) history.back() = (double)(true_norm / ref);
                    return SFEM_SUCCESS;
                }
                rnorm = true_norm;
                rtzold = blas->dot(n, r, z);  // r was replaced by the true residual
            }
            precondition(n, r, z);
            const T rtz_new = blas->dot(n, r, z);
            const T beta = (rtz_new - rtzold) / rtz;
            rtz = rtz_new;
            blas->axpby(n, T(1), z, beta, p);
        }
        return SFEM_FAILURE;
    }
};

}  // namespace sfem
