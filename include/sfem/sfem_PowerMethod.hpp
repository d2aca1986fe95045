// sfem::PowerMethod<T> (algebra/sfem_PowerMethod.hpp:11-56): largest eigenvalue of A by normalised
// power iteration.  The start vector's product is normalised into `work`; the iteration then
// alternates its destination between `eigen_vector` and `work` and stops on a change below tol only
// when the newest iterate sits in `eigen_vector` (at most max_it iterations, rounded up to odd),
// which is the reference's iteration sequence step for step.
#pragma once

#include <cmath>
#include <utility>

#include "sfem_tpl_blas.hpp"

namespace sfem {

template <typename T>
class PowerMethod {
public:
    std::shared_ptr<BLAS<T>> blas;
    int last_iterations{0};

    T max_eigen_value(const std::function<void(const T *const, T *const)> &op, int max_it, T tol, std::ptrdiff_t n,
                      T *eigen_vector, T *work) {
        // dst = A src / |A src|; returns |A src|
        auto normalised_product = [&](const T *src, T *dst) {
            blas->zeros(n, dst);
            op(src, dst);
            const T norm = blas->norm2(n, dst);
            blas->scal(n, T(1) / norm, dst);
            return norm;
        };
        normalised_product(eigen_vector, work);
        const int limit = (max_it % 2 == 0) ? max_it + 1 : max_it;
        T *dst = eigen_vector, *src = work;
        T estimate = 0, previous = 0;
        for (last_iterations = 1; last_iterations <= limit; last_iterations++) {
            estimate = normalised_product(src, dst);
            const bool in_place = dst == eigen_vector;
            if (in_place && std::abs(estimate - previous) <= tol) break;
            previous = estimate;
            std::swap(dst, src);
        }
        if (last_iterations > limit) last_iterations = limit;
        return estimate;
    }
};

}  // namespace sfem
