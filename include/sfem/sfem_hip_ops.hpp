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

}  // namespace sfem
