// sfem::Operator<T> and helpers: the interface of algebra/sfem_Operator.hpp:12-188.
#pragma once

#include <cassert>

#include "sfem_base.hpp"

namespace sfem {

template <typename T>
class Operator {
public:
    virtual ~Operator() = default;
    // y += A x for operators; x := solve(b) for solvers (x is the initial guess)
    virtual int apply(const T *const x, T *const y) = 0;
    virtual int multi_apply(const int vec_size, const T *const *const x, T *const *const y) {
        for (int b = 0; b < vec_size; b++) {
            if (apply(x[b], y[b]) != SFEM_SUCCESS) return SFEM_FAILURE;
        }
        return SFEM_SUCCESS;
    }
    virtual std::ptrdiff_t rows() const = 0;
    virtual std::ptrdiff_t cols() const = 0;
    virtual ExecutionSpace execution_space() const = 0;
};

template <typename T>
using SharedOperator = std::shared_ptr<Operator<T>>;

// y := f(y) on one vector (algebra/sfem_Operator.hpp:35-42): constraints, scalings, projections
template <typename T>
class InPlaceOperator {
public:
    virtual ~InPlaceOperator() = default;
    virtual int apply(T *const y) = 0;
    virtual std::ptrdiff_t size() const = 0;
    virtual ExecutionSpace execution_space() const = 0;
};

template <typename T>
using SharedInPlaceOperator = std::shared_ptr<InPlaceOperator<T>>;

template <typename T>
class LambdaInPlaceOperator final : public InPlaceOperator<T> {
public:
    std::ptrdiff_t size_{0};
    std::function<void(T *const)> apply_;
    ExecutionSpace execution_space_;
    LambdaInPlaceOperator(std::ptrdiff_t n, std::function<void(T *const)> f, ExecutionSpace es)
        : size_(n), apply_(std::move(f)), execution_space_(es) {}
    int apply(T *const y) override {
        apply_(y);
        return SFEM_SUCCESS;
    }
    std::ptrdiff_t size() const override { return size_; }
    ExecutionSpace execution_space() const override { return execution_space_; }
};

template <typename T>
inline SharedInPlaceOperator<T> make_in_place_op(std::ptrdiff_t n, std::function<void(T *const)> f,
                                                 ExecutionSpace es) {
    return std::make_shared<LambdaInPlaceOperator<T>>(n, std::move(f), es);
}

template <typename T>
class LambdaOperator final : public Operator<T> {
public:
    std::ptrdiff_t rows_{0}, cols_{0};
    std::function<void(const T *const, T *const)> apply_;
    ExecutionSpace execution_space_;
    LambdaOperator(std::ptrdiff_t rows, std::ptrdiff_t cols, std::function<void(const T *const, T *const)> f,
                   ExecutionSpace es)
        : rows_(rows), cols_(cols), apply_(std::move(f)), execution_space_(es) {}
    int apply(const T *const x, T *const y) override {
        apply_(x, y);
        return SFEM_SUCCESS;
    }
    std::ptrdiff_t rows() const override { return rows_; }
    std::ptrdiff_t cols() const override { return cols_; }
    ExecutionSpace execution_space() const override { return execution_space_; }
};

template <typename T>
inline SharedOperator<T> make_op(std::ptrdiff_t rows, std::ptrdiff_t cols,
                                 std::function<void(const T *const, T *const)> op, ExecutionSpace es) {
    return std::make_shared<LambdaOperator<T>>(rows, cols, std::move(op), es);
}

// (A + B) x: both accumulate into y
template <typename T>
class OperatorAdd final : public Operator<T> {
public:
    OperatorAdd(SharedOperator<T> l, SharedOperator<T> r) : l_(std::move(l)), r_(std::move(r)) {
        assert(l_->rows() == r_->rows() && l_->cols() == r_->cols());
    }
    int apply(const T *const x, T *const y) override { return r_->apply(x, y) | l_->apply(x, y); }
    std::ptrdiff_t rows() const override { return l_->rows(); }
    std::ptrdiff_t cols() const override { return l_->cols(); }
    ExecutionSpace execution_space() const override { return l_->execution_space(); }

private:
    SharedOperator<T> l_, r_;
};

// Composition with the reference's semantics (algebra/sfem_Operator.hpp:77-99)
template <typename T>
class OperatorCompose final : public Operator<T> {
public:
    OperatorCompose(SharedOperator<T> l, SharedOperator<T> r) : l_(std::move(l)), r_(std::move(r)) {}
    int apply(const T *const x, T *const y) override { return r_->apply(x, y) | l_->apply(x, y); }
    std::ptrdiff_t rows() const override { return l_->rows(); }
    std::ptrdiff_t cols() const override { return r_->cols(); }
    ExecutionSpace execution_space() const override { return l_->execution_space(); }

private:
    SharedOperator<T> l_, r_;
};

// An in-place operator after an operator (algebra/sfem_Operator.hpp:99-117): r applied to (x, y),
// then l to y
template <typename T>
class InPlaceAfterOperator final : public Operator<T> {
public:
    InPlaceAfterOperator(SharedInPlaceOperator<T> l, SharedOperator<T> r) : l_(std::move(l)), r_(std::move(r)) {
        assert(l_->size() == r_->rows());
    }
    int apply(const T *const x, T *const y) override { return r_->apply(x, y) | l_->apply(y); }
    std::ptrdiff_t rows() const override { return l_->size(); }
    std::ptrdiff_t cols() const override { return r_->cols(); }
    ExecutionSpace execution_space() const override { return l_->execution_space(); }

private:
    SharedInPlaceOperator<T> l_;
    SharedOperator<T> r_;
};

template <typename T>
inline SharedOperator<T> operator*(const SharedInPlaceOperator<T> &l, const SharedOperator<T> &r) {
    return std::make_shared<InPlaceAfterOperator<T>>(l, r);
}

template <typename T>
inline SharedOperator<T> operator+(const SharedOperator<T> &l, const SharedOperator<T> &r) {
    return std::make_shared<OperatorAdd<T>>(l, r);
}
template <typename T>
inline SharedOperator<T> operator*(const SharedOperator<T> &l, const SharedOperator<T> &r) {
    return std::make_shared<OperatorCompose<T>>(l, r);
}

}  // namespace sfem
