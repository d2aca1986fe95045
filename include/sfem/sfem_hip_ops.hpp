// sfem::Operator<T> views of the library's device operators (sfem_amd_op handles):
// FE operators, transfers and the C-ABI solvers, usable in the sfem:: solver templates.
#pragma once

#include <memory>
#include <utility>

#include "sfem_Operator.hpp"

namespace sfem {

template <typename T>
class DeviceOperator final : public Operator<T> {
public:
    // Takes shared ownership of `op`: destroyed with the last reference.  `keep` (optional) is held
    // until then as well: e.g. the meshes the operator was created on.
    explicit DeviceOperator(sfem_amd_op *op, void *stream = nullptr, std::shared_ptr<void> keep = nullptr)
        : keep_(std::move(keep)), op_(op, [](sfem_amd_op *p) { sfem_amd_op_destroy(p); }), stream_(stream) {}
    int apply(const T *const x, T *const y) override { return sfem_amd_op_apply(op_.get(), x, y, stream_); }
    std::ptrdiff_t rows() const override { return sfem_amd_op_rows(op_.get()); }
    std::ptrdiff_t cols() const override { return sfem_amd_op_cols(op_.get()); }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }
    sfem_amd_op *handle() const { return op_.get(); }

private:
    std::shared_ptr<void> keep_;  // declared first: released after op_
    std::shared_ptr<sfem_amd_op> op_;
    void *stream_;
};

template <typename T>
inline SharedOperator<T> make_device_op(sfem_amd_op *op, void *stream = nullptr, std::shared_ptr<void> keep = nullptr) {
    return std::make_shared<DeviceOperator<T>>(op, stream, std::move(keep));
}

// y[c] = value on the constrained dofs c of a constrained operator (sfem_amd_constrained_op_create),
// in place: constraint_nodes_to_value_vec (operators/boundary_conditions/dirichlet.cpp:95-106) as an
// InPlaceOperator, e.g. to zero the Dirichlet rows of a right-hand side on the device
template <typename T>
class DeviceConstraintValue final : public InPlaceOperator<T> {
public:
    DeviceConstraintValue(std::shared_ptr<DeviceOperator<T>> constrained, double value, void *stream = nullptr)
        : c_(std::move(constrained)), value_(value), stream_(stream) {}
    int apply(T *const y) override { return sfem_amd_constrained_op_set_value(c_->handle(), value_, y, stream_); }
    std::ptrdiff_t size() const override { return c_->rows(); }
    ExecutionSpace execution_space() const override { return EXECUTION_SPACE_DEVICE; }

private:
    std::shared_ptr<DeviceOperator<T>> c_;
    double value_;
    void *stream_;
};

// `constrained` must view a constrained operator handle (a make_device_op of sfem_amd_constrained_op_create)
template <typename T>
inline SharedInPlaceOperator<T> make_constraint_value_op(const SharedOperator<T> &constrained, double value,
                                                         void *stream = nullptr) {
    auto d = std::dynamic_pointer_cast<DeviceOperator<T>>(constrained);
    if (!d) fatal("make_constraint_value_op: not a device operator");
    return std::make_shared<DeviceConstraintValue<T>>(std::move(d), value, stream);
}

}  // namespace sfem
