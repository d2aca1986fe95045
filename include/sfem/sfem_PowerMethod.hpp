// sfem::PowerMethod<T> (algebra/sfem_PowerMethod.hpp:11-56): largest eigenvalue by power
// iteration; the iteration count is made odd so the result lands in `eigen_vector`.
#pragma once

#include <cmath>

#include "sfem_tpl_blas.hpp"

namespace sfem {

template <typename T>
class PowerMethod {
public:
    std::shared_ptr<BLAS<T>> blas;
    int last_iterations{0};

    T max_eigen_value(const std::function<void(const T *const, T *const)> &op, int max_it, T tol, std::ptrdiff_t n,
                      T *eigen_vector, T *work) {
        const int max_it_odd = max_it + (max_it % 2 == 0);
        blas->zeros(n, work);
        op(eigen_vector, work);
        blas->scal(n, T(1) / blas->norm2(n, work), work);
        T *vecs[2] = {eigen_vector, work};
        T lambda = 0, lambda_old = 0;
        last_iterations = 0;
        for (int k = 0; k < max_it_odd; k++) {
            T *w = vecs[k % 2];
            T *ev = vecs[(k + 1) % 2];
            blas->zeros(n, w);
            op(ev, w);
            lambda = blas->norm2(n, w);
            blas->scal(n, T(1) / lambda, w);
            last_iterations = k + 1;
            if (std::abs(lambda - lambda_old) <= tol && k % 2 == 0) break;
            lambda_old = lambda;
        }
        return lambda;
    }
};

}  // namespace sfem
