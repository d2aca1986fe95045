// sfem::Multigrid<T> (algebra/sfem_Multigrid.hpp:27-347).  Level 0 is the finest;
// add_level(op, smoother_or_solver, prolongation, restriction).  apply(rhs, x) wraps the
// caller's arrays as finest-level storage and runs up to max_it cycles, stopping when the
// finest residual norm is below atol.
#pragma once

#include <cmath>
#include <vector>

#include "sfem_Buffer.hpp"
#include "sfem_MatrixFreeLinearSolver.hpp"
#include "sfem_Operator.hpp"
#include "sfem_tpl_blas.hpp"

namespace sfem {

// Extension: a prolongation that adds its result into the fine iterate (y += P x, constrained fine
// rows unchanged) — the coarse-grid correction in one pass instead of zero + apply + axpby.
template <typename T>
class AddingProlongation {
public:
    virtual ~AddingProlongation() = default;
    virtual bool adds_correction() = 0;
    virtual int apply_add(const T *coarse, T *fine) = 0;
};

template <typename T>
class Multigrid final : public Operator<T> {
public:
    enum CycleType { V_CYCLE = 1, W_CYCLE = 2 };
    enum CycleReturnCode { CYCLE_CONTINUE = 0, CYCLE_CONVERGED = 1, CYCLE_FAILURE = 2 };

    std::shared_ptr<BLAS<T>> blas_;
    bool verbose{false};
    bool debug{false};  // with verbose: the residual norm of every level's cycle, not only the finest
    std::vector<double> residual_history;  // finest residual norm per cycle
    // Called after each cycle with (cycle, finest residual norm); may replace smoothers.
    std::function<void(int, double)> on_cycle;

    void default_init(void *stream = nullptr) { blas_ = make_device_blas<T>(stream); }
    std::shared_ptr<BLAS<T>> &blas() { return blas_; }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }

    void add_level(const SharedOperator<T> &op, const SharedOperator<T> &smoother_or_solver,
                   const SharedOperator<T> &prolongation, const SharedOperator<T> &restriction) {
        operator_.push_back(op);
        smoother_.push_back(smoother_or_solver);
        prolongation_.push_back(prolongation);
        restriction_.push_back(restriction);
        memory_.clear();
    }
    void set_cycle_type(int v) { cycle_type_ = v; }
    // x is zero on entry to apply() (e.g. a preconditioner whose caller zeroes its output): the
    // first finest-level pre-smoothing then skips its product with the zero iterate.
    void set_initial_guess_zero(bool v) { is_initial_guess_zero_ = v; }
    void set_max_it(int v) { max_it_ = v; }
    void set_atol(T v) { atol_ = v; }
    int n_levels() const { return (int)smoother_.size(); }
    int iterations() const { return iterations_; }
    std::ptrdiff_t rows() const override { return operator_.front()->rows(); }
    std::ptrdiff_t cols() const override { return operator_.front()->cols(); }
    std::vector<SharedOperator<T>> &operators() { return operator_; }
    std::vector<SharedOperator<T>> &smoothers() { return smoother_; }

    int apply(const T *const rhs, T *const x) override {
        ensure_init();
        memory_[0].solution = Buffer<T>::wrap(rows(), x);
        memory_[0].rhs = Buffer<T>::wrap(rows(), (T *)rhs);
        residual_history.clear();
        for (iterations_ = 0; iterations_ < max_it_; iterations_++) {
            const CycleReturnCode ret = cycle(0);
            if (ret == CYCLE_FAILURE) return SFEM_FAILURE;
            if (ret == CYCLE_CONVERGED) break;
        }
        return SFEM_SUCCESS;
    }

private:
    struct Memory {
        SharedBuffer<T> rhs, solution, work;
        std::ptrdiff_t size() const { return work->size(); }
    };
    std::vector<SharedOperator<T>> operator_, smoother_, prolongation_, restriction_;
    std::vector<Memory> memory_;
public:
    // the level operators overwrite their output (as ConjugateGradient::apply_overwrites_output)
    bool apply_overwrites_output{false};
    // the restrictions overwrite their output (no zero fill of the coarse right-hand side)
    bool restriction_overwrites_output{false};

private:
    int max_it_{10};
    int iterations_{0};
    bool is_initial_guess_zero_{false};
    int cycle_type_{V_CYCLE};
    T atol_{1e-10};

    void ensure_init() {
        if (!memory_.empty()) return;
        memory_.resize(n_levels());
        auto b = blas_;
        for (int l = 0; l < n_levels(); l++) {
            const std::ptrdiff_t n = operator_[l]->rows();
            if (l != 0) {
                memory_[l].solution = Buffer<T>::own(n, b->allocate(n), [b](void *q) { b->destroy(q); });
                memory_[l].rhs = Buffer<T>::own(n, b->allocate(n), [b](void *q) { b->destroy(q); });
            }
            memory_[l].work = Buffer<T>::own(n, b->allocate(n), [b](void *q) { b->destroy(q); });
        }
    }

    // Runs a smoother that may be a MatrixFreeLinearSolver in its zero-initial-guess mode; smoothers
    // inside a multigrid otherwise run with a general initial guess (post-smoothing needs it).
    static void pre_smooth(const SharedOperator<T> &smoother, bool zero_x, const T *rhs, T *x) {
        auto *mf = zero_x ? dynamic_cast<MatrixFreeLinearSolver<T> *>(smoother.get()) : nullptr;
        if (mf) mf->set_initial_guess_zero(true);
        smoother->apply(rhs, x);
        if (mf) mf->set_initial_guess_zero(false);
    }

    // algebra/sfem_Multigrid.hpp:242-346
    CycleReturnCode cycle(int level) {
        Memory &mem = memory_[level];
        auto &smoother = smoother_[level];
        if (level == n_levels() - 1) {
            blas_->zeros(mem.solution->size(), mem.solution->data());
            return smoother->apply(mem.rhs->data(), mem.solution->data()) == SFEM_SUCCESS ? CYCLE_CONTINUE
                                                                                         : CYCLE_FAILURE;
        }
        auto &op = operator_[level];
        auto &restriction = restriction_[level];
        auto &prolongation = prolongation_[level + 1];
        Memory &coarse = memory_[level + 1];
        for (int k = 0; k < cycle_type_; k++) {
            // coarse solutions arrive zeroed (below), the finest one when the caller says so: the
            // smoother's first residual is then rhs itself, and A 0 is not computed (same result)
            const bool zero_x = k == 0 && (level > 0 || (is_initial_guess_zero_ && iterations_ == 0));
            pre_smooth(smoother, zero_x, mem.rhs->data(), mem.solution->data());
            // residual
            if (!apply_overwrites_output) blas_->zeros(mem.size(), mem.work->data());
            op->apply(mem.solution->data(), mem.work->data());
            blas_->axpby(mem.size(), T(1), mem.rhs->data(), T(-1), mem.work->data());
            // the finest residual norm (a host synchronisation) is only needed to stop or report;
            // a single cycle used as a preconditioner skips it, so the cycle can be graph-captured
            if (level == 0 && (max_it_ > 1 || on_cycle || verbose || atol_ > 0)) {
                const double nr = (double)blas_->norm2(mem.size(), mem.work->data());
                residual_history.push_back(nr);
                if (verbose) std::printf("Multigrid %d\t%g\n", iterations_, nr);
                if (on_cycle) on_cycle(iterations_, nr);
                if (nr < (double)atol_) return CYCLE_CONVERGED;
            } else if (level > 0 && verbose && debug) {
                std::printf("  level %d residual %g\n", level, (double)blas_->norm2(mem.size(), mem.work->data()));
            }
            // restriction
            if (!restriction_overwrites_output) blas_->zeros(coarse.rhs->size(), coarse.rhs->data());
            restriction->apply(mem.work->data(), coarse.rhs->data());
            // the coarse solution starts at zero: the coarsest level zeroes it itself, and a smoother
            // that does not read x in its zero-guess mode needs no fill either
            const bool coarsest_next = level + 1 == n_levels() - 1;
            auto *csm = coarsest_next ? nullptr : dynamic_cast<MatrixFreeLinearSolver<T> *>(smoother_[level + 1].get());
            if (!coarsest_next && !(csm && csm->zero_guess_ignores_x()))
                blas_->zeros(coarse.solution->size(), coarse.solution->data());
            if (cycle(level + 1) == CYCLE_FAILURE) return CYCLE_FAILURE;
            // prolongation and correction
            auto *ap = dynamic_cast<AddingProlongation<T> *>(prolongation.get());
            if (ap && ap->adds_correction()) {
                if (ap->apply_add(coarse.solution->data(), mem.solution->data())) return CYCLE_FAILURE;
            } else {
                blas_->zeros(mem.size(), mem.work->data());
                prolongation->apply(coarse.solution->data(), mem.work->data());
                blas_->axpby(mem.size(), T(1), mem.work->data(), T(1), mem.solution->data());
            }
            smoother->apply(mem.rhs->data(), mem.solution->data());
        }
        return CYCLE_CONTINUE;
    }
};

}  // namespace sfem
