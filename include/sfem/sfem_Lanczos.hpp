// Extension (the reference's Chebyshev3 takes its bound from PowerMethod only): the largest
// eigenvalue of a symmetric positive operator A, or of D^-1 A for a positive diagonal D, by the
// Lanczos process.  D^-1 A is self-adjoint in the inner product <u, v>_D = u^T D v, so the process
// runs in that inner product; with u = D v kept beside v every product it needs is a plain dot:
//   y = A v_j,  alpha_j = v_j^T y,  D w = y - alpha_j u_j - beta_{j-1} u_{j-1},  w = D^-1 (D w),
//   beta_j = sqrt(w^T D w),  v_{j+1} = w / beta_j,  u_{j+1} = D w / beta_j.
// The estimate is the largest eigenvalue of the tridiagonal matrix (alpha, beta) (Sturm
// bisection), a Ritz value that approaches lambda_max from below far faster than power
// iteration; the process stops when it changes by at most tol relative, when beta vanishes (an
// invariant subspace) or after max_it steps.  No reorthogonalisation: its loss only adds copies
// of converged Ritz values, which leaves the largest one unchanged.
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <vector>

#include "sfem_tpl_blas.hpp"

namespace sfem {

// Largest eigenvalue of the symmetric tridiagonal matrix with diagonal a and off-diagonal b
// (b.size() == a.size() - 1): bisection on the Gershgorin interval with the Sturm count of the
// LDL^T pivots of T - s I (the number of negative pivots is the number of eigenvalues below s).
inline double tridiagonal_max_eigenvalue(const std::vector<double> &a, const std::vector<double> &b) {
    const int k = (int)a.size();
    if (k == 0) return 0;
    double lo = a[0], hi = a[0];
    for (int i = 0; i < k; i++) {
        const double r = (i > 0 ? std::abs(b[i - 1]) : 0.0) + (i + 1 < k ? std::abs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    const double tiny = 1e-300;
    auto count_below = [&](double s) {
        int c = 0;
        double d = 1;
        for (int i = 0; i < k; i++) {
            d = a[i] - s - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
            if (std::abs(d) < tiny) d = -tiny;
            if (d < 0) c++;
        }
        return c;
    };
    for (int it = 0; it < 200; it++) {
        const double m = 0.5 * (lo + hi);
        if (m <= lo || m >= hi) break;
        if (count_below(m) == k)
            hi = m;  // every eigenvalue lies below m
        else
            lo = m;
    }
    return hi;
}

template <typename T>
class Lanczos {
public:
    std::shared_ptr<BLAS<T>> blas;
    int last_iterations{0};

    // op(x, y): y = A x into a zeroed y.  inv_diag: D^-1 (device), or null for A itself.  start:
    // the first vector of the process in D-scaled form (u_1 = start, v_1 = D^-1 start); dofs that
    // are zero in it and kept zero by A (constrained rows) stay out of the estimate.
    T max_eigen_value(const std::function<void(const T *const, T *const)> &op, const T *inv_diag, int max_it, T tol,
                      std::ptrdiff_t n, const T *start) {
        const bool scaled = inv_diag != nullptr;
        T *u = blas->allocate(n), *u_prev = blas->allocate(n), *y = blas->allocate(n);
        T *v = scaled ? blas->allocate(n) : u;
        blas->copy(n, start, u);
        blas->zeros(n, u_prev);
        if (scaled) blas->xypaz(n, inv_diag, u, T(0), v);
        std::vector<double> a, b;
        double estimate = 0, beta = std::sqrt(std::max(0.0, (double)blas->dot(n, v, u)));
        last_iterations = 0;
        if (beta > 0) {
            blas->scal(n, T(1 / beta), u);
            if (scaled) blas->scal(n, T(1 / beta), v);
            beta = 0;
            for (last_iterations = 1; last_iterations <= std::max(1, max_it); last_iterations++) {
                blas->zeros(n, y);
                op(v, y);
                const double alpha = (double)blas->dot(n, v, y);
                a.push_back(alpha);
                // D w = y - alpha u_j - beta u_{j-1}
                blas->axpy(n, T(-alpha), u, y);
                if (beta != 0) blas->axpy(n, T(-beta), u_prev, y);
                const double previous = estimate;
                estimate = tridiagonal_max_eigenvalue(a, b);
                // w = D^-1 (D w) into v (v_j is no longer needed), then beta_j = sqrt(w^T D w)
                if (scaled) blas->xypaz(n, inv_diag, y, T(0), v);
                const double wDw = (double)blas->dot(n, scaled ? v : y, y);
                const double beta_new = std::sqrt(std::max(0.0, wDw));
                if (last_iterations > 1 && std::abs(estimate - previous) <= (double)tol * std::abs(estimate)) break;
                if (!(beta_new > 1e-14 * std::abs(estimate)) || last_iterations == std::max(1, max_it)) break;
                b.push_back(beta_new);
                // rotate: u_{j-1} <- u_j, u_{j+1} = D w / beta, v_{j+1} = w / beta
                std::swap(u_prev, u);
                std::swap(u, y);
                blas->scal(n, T(1 / beta_new), u);
                if (scaled)
                    blas->scal(n, T(1 / beta_new), v);
                else
                    v = u;
                beta = beta_new;
            }
        }
        blas->destroy(u);
        blas->destroy(u_prev);
        blas->destroy(y);
        if (scaled) blas->destroy(v);
        return (T)estimate;
    }
};

}  // namespace sfem
