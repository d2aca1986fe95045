// Host-side semantics of the sfem:: operator algebra (include/sfem/sfem_Operator.hpp) on lambda
// operators, the composition rules of algebra/sfem_Operator.hpp:44-117: OperatorAdd and
// OperatorCompose apply the right operand, then the left one, both to (x, y); an in-place operator
// composed after an operator applies to y.  Prints "ok" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sfem/sfem_Operator.hpp"

using sfem::SharedOperator;

int main() {
    const std::ptrdiff_t n = 4;
    const auto host = sfem::EXECUTION_SPACE_HOST;
    // A x = 2 x (accumulating), B x = x + 1 (accumulating)
    SharedOperator<double> A = sfem::make_op<double>(
        n, n, [=](const double *x, double *y) { for (std::ptrdiff_t i = 0; i < n; i++) y[i] += 2 * x[i]; }, host);
    SharedOperator<double> B = sfem::make_op<double>(
        n, n, [=](const double *x, double *y) { for (std::ptrdiff_t i = 0; i < n; i++) y[i] += x[i] + 1; }, host);
    // C y = -y (in place)
    sfem::SharedInPlaceOperator<double> C = sfem::make_in_place_op<double>(
        n, [=](double *y) { for (std::ptrdiff_t i = 0; i < n; i++) y[i] = -y[i]; }, host);
    const std::vector<double> x{1, 2, 3, 4};
    int bad = 0;
    auto check = [&](const char *what, const std::vector<double> &got, const std::vector<double> &want) {
        for (std::ptrdiff_t i = 0; i < n; i++)
            if (std::fabs(got[i] - want[i]) > 1e-15) {
                std::printf("%s: entry %td is %g, expected %g\n", what, i, got[i], want[i]);
                bad = 1;
                return;
            }
    };
    std::vector<double> y(n, 0.5);
    (A + B)->apply(x.data(), y.data());  // 0.5 + (x + 1) + 2x
    check("A + B", y, {4.5, 7.5, 10.5, 13.5});
    y.assign(n, 0);
    (A * B)->apply(x.data(), y.data());  // both accumulate into y: B then A
    check("A * B", y, {4, 7, 10, 13});
    y.assign(n, 1);
    (C * A)->apply(x.data(), y.data());  // -(1 + 2x)
    check("C * A", y, {-3, -5, -7, -9});
    if ((C * A)->rows() != n || (C * A)->cols() != n || C->size() != n) bad = 1;
    std::printf(bad ? "failed\n" : "ok\n");
    return bad;
}
