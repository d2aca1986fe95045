// Host checks of the Lanczos eigenvalue bound (include/sfem/sfem_Lanczos.hpp) against closed forms:
//  * the tridiagonal bisection on T = tridiag(-1, 2, -1) of order n, lambda_max = 2 + 2 cos(pi/(n+1));
//  * Lanczos on the 1-D Laplacian (same closed form), to full accuracy when run to convergence and
//    from below (a Ritz value) when cut short;
//  * Lanczos on D^-1 A with A = D^1/2 T D^1/2 for a positive diagonal D: D^-1 A is similar to T;
//  * a start vector that is zero on rows the operator keeps to themselves (identity rows scaled
//    by 10, above the spectrum) leaves them out of the estimate.
// Prints "ok" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "sfem/sfem_Lanczos.hpp"

namespace {

// a host BLAS on new[] arrays (what the Lanczos process calls)
class HostBLAS final : public sfem::BLAS<double> {
public:
    double *allocate(std::size_t n) override { return new double[n > 0 ? n : 1](); }
    void destroy(void *a) override { delete[] (double *)a; }
    void zeros(std::size_t n, double *x) override { std::memset(x, 0, n * sizeof(double)); }
    void values(std::size_t n, double v, double *x) override {
        for (std::size_t i = 0; i < n; i++) x[i] = v;
    }
    void copy(std::ptrdiff_t n, const double *s, double *d) override { std::memcpy(d, s, n * sizeof(double)); }
    double dot(std::ptrdiff_t n, const double *l, const double *r) override {
        double s = 0;
        for (std::ptrdiff_t i = 0; i < n; i++) s += l[i] * r[i];
        return s;
    }
    void axpy(std::ptrdiff_t n, double a, const double *x, double *y) override {
        for (std::ptrdiff_t i = 0; i < n; i++) y[i] += a * x[i];
    }
    void axpby(std::ptrdiff_t n, double a, const double *x, double b, double *y) override {
        for (std::ptrdiff_t i = 0; i < n; i++) y[i] = a * x[i] + b * y[i];
    }
    void scal(std::ptrdiff_t n, double a, double *x) override {
        for (std::ptrdiff_t i = 0; i < n; i++) x[i] *= a;
    }
    void reciprocal(std::ptrdiff_t n, double a, double *x) override {
        for (std::ptrdiff_t i = 0; i < n; i++) x[i] = a / x[i];
    }
    double norm2(std::ptrdiff_t n, const double *x) override { return std::sqrt(dot(n, x, x)); }
    void zaxpby(std::ptrdiff_t n, double a, const double *x, double b, const double *y, double *z) override {
        for (std::ptrdiff_t i = 0; i < n; i++) z[i] = a * x[i] + b * y[i];
    }
    void xypaz(std::ptrdiff_t n, const double *x, const double *y, double a, double *z) override {
        for (std::ptrdiff_t i = 0; i < n; i++) z[i] = x[i] * y[i] + a * z[i];
    }
};

double laplacian_1d_max(int n) { return 2 + 2 * std::cos(M_PI / (n + 1)); }

}  // namespace

int main() {
    int bad = 0;
    auto expect = [&](const char *what, double got, double want, double tol) {
        if (!(std::abs(got - want) <= tol * std::abs(want))) {
            std::printf("%s: %.17g, expected %.17g (tol %g)\n", what, got, want, tol);
            bad = 1;
        }
    };
    // tridiagonal bisection
    for (int n : {1, 2, 7, 64}) {
        std::vector<double> a(n, 2.0), b(n > 0 ? n - 1 : 0, -1.0);
        expect("tridiagonal bisection", sfem::tridiagonal_max_eigenvalue(a, b), laplacian_1d_max(n), 1e-14);
    }
    auto blas = std::make_shared<HostBLAS>();
    const int n = 200;
    auto lap = [n](const double *x, double *y) {
        for (int i = 0; i < n; i++) y[i] += 2 * x[i] - (i > 0 ? x[i - 1] : 0) - (i + 1 < n ? x[i + 1] : 0);
    };
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.5, 1.5);
    std::vector<double> start(n);
    for (auto &s : start) s = U(rng);
    sfem::Lanczos<double> lz;
    lz.blas = blas;
    // to convergence
    const double full = lz.max_eigen_value(lap, nullptr, 400, 1e-15, n, start.data());
    expect("Lanczos, 1-D Laplacian", full, laplacian_1d_max(n), 1e-12);
    // cut short: a Ritz value, below lambda_max and close to it
    const double short10 = lz.max_eigen_value(lap, nullptr, 10, 0, n, start.data());
    if (!(short10 <= laplacian_1d_max(n) * (1 + 1e-14) && short10 > 0.9 * laplacian_1d_max(n))) {
        std::printf("Lanczos, 10 steps: %.17g not a lower bound close to %.17g\n", short10, laplacian_1d_max(n));
        bad = 1;
    }
    if (lz.last_iterations != 10) {
        std::printf("Lanczos, 10 steps: ran %d\n", lz.last_iterations);
        bad = 1;
    }
    // D^-1 A with A = D^1/2 T D^1/2
    std::vector<double> d(n), sq(n), inv(n);
    for (int i = 0; i < n; i++) {
        d[i] = std::pow(10.0, 3 * (U(rng) - 0.5));  // three decades
        sq[i] = std::sqrt(d[i]);
        inv[i] = 1 / d[i];
    }
    auto scaled = [&](const double *x, double *y) {
        std::vector<double> t(n), u(n, 0.0);
        for (int i = 0; i < n; i++) t[i] = sq[i] * x[i];
        lap(t.data(), u.data());
        for (int i = 0; i < n; i++) y[i] += sq[i] * u[i];
    };
    const double scaled_full = lz.max_eigen_value(scaled, inv.data(), 400, 1e-15, n, start.data());
    expect("Lanczos, D^-1 A", scaled_full, laplacian_1d_max(n), 1e-11);
    // rows kept to themselves (y_i = 10 x_i, no coupling) with a zero start entry stay out
    const int m = 50;
    auto constrained = [m](const double *x, double *y) {
        for (int i = 0; i < m; i++) {
            if (i == 0 || i == m - 1) {
                y[i] += 10 * x[i];
                continue;
            }
            y[i] += 2 * x[i] - (i > 1 ? x[i - 1] : 0) - (i + 2 < m ? x[i + 1] : 0);
        }
    };
    // (a random interior: a symmetric start such as ones is orthogonal to the antisymmetric top
    // mode of this symmetric chain and would find the next mode, 2 + 2 cos(2 pi / (m - 1)))
    std::vector<double> cstart(m);
    for (auto &s : cstart) s = U(rng);
    cstart[0] = cstart[m - 1] = 0;
    const double cfull = lz.max_eigen_value(constrained, nullptr, 200, 1e-15, m, cstart.data());
    expect("Lanczos, zero start on decoupled rows", cfull, laplacian_1d_max(m - 2), 1e-12);
    std::printf(bad ? "failed\n" : "ok\n");
    return bad;
}
