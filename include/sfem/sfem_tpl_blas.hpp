// sfem::BLAS<T> (algebra/sfem_tpl_blas.hpp:11-35) and its device implementation over the
// C ABI (replaces algebra/cuda/sfem_cuda_blas.hpp:55-73; reductions are the library's own
// deterministic kernels, no vendor BLAS).
#pragma once

#include "sfem_base.hpp"

namespace sfem {

template <typename T>
class BLAS {
public:
    virtual ~BLAS() = default;
    virtual T *allocate(std::size_t n) = 0;
    virtual void destroy(void *a) = 0;
    virtual void zeros(std::size_t n, T *x) = 0;
    virtual void values(std::size_t n, T value, T *x) = 0;
    virtual void copy(std::ptrdiff_t n, const T *src, T *dest) = 0;
    virtual T dot(std::ptrdiff_t n, const T *l, const T *r) = 0;
    virtual void axpy(std::ptrdiff_t n, T alpha, const T *x, T *y) = 0;
    virtual void axpby(std::ptrdiff_t n, T alpha, const T *x, T beta, T *y) = 0;
    virtual void scal(std::ptrdiff_t n, T alpha, T *x) = 0;
    virtual void reciprocal(std::ptrdiff_t n, T alpha, T *x) = 0;
    virtual T norm2(std::ptrdiff_t n, const T *x) = 0;
    virtual void zaxpby(std::ptrdiff_t n, T alpha, const T *x, T beta, const T *y, T *z) = 0;
    // z = x * y + alpha * z
    virtual void xypaz(std::ptrdiff_t n, const T *x, const T *y, T alpha, T *z) = 0;
    virtual bool good() const { return true; }
};

template <typename T>
class DeviceBLAS final : public BLAS<T> {
    static constexpr int rt = real_type_of<T>::value;

public:
    explicit DeviceBLAS(void *stream = nullptr) : s_(stream) {}
    T *allocate(std::size_t n) override {
        T *p = (T *)sfem_amd_malloc(sizeof(T) * (n > 0 ? n : 1));
        if (!p) fatal("DeviceBLAS::allocate");
        sfem_amd_memset(p, 0, sizeof(T) * n, s_);
        return p;
    }
    void destroy(void *a) override { sfem_amd_free(a); }
    // a fill kernel, not a memset: solver sequences may be captured into HIP graphs
    void zeros(std::size_t n, T *x) override { sfem_amd_blas_values(rt, (std::ptrdiff_t)n, 0.0, x, s_); }
    void values(std::size_t n, T v, T *x) override { sfem_amd_blas_values(rt, (std::ptrdiff_t)n, (double)v, x, s_); }
    void copy(std::ptrdiff_t n, const T *src, T *dst) override { sfem_amd_blas_copy(rt, n, src, dst, s_); }
    T dot(std::ptrdiff_t n, const T *l, const T *r) override { return (T)sfem_amd_blas_dot(rt, n, l, r, s_); }
    void axpy(std::ptrdiff_t n, T a, const T *x, T *y) override { sfem_amd_blas_axpby(rt, n, a, x, 1.0, y, s_); }
    void axpby(std::ptrdiff_t n, T a, const T *x, T b, T *y) override { sfem_amd_blas_axpby(rt, n, a, x, b, y, s_); }
    void scal(std::ptrdiff_t n, T a, T *x) override { sfem_amd_blas_scal(rt, n, a, x, s_); }
    void reciprocal(std::ptrdiff_t n, T a, T *x) override { sfem_amd_blas_reciprocal(rt, n, a, x, s_); }
    T norm2(std::ptrdiff_t n, const T *x) override { return (T)sfem_amd_blas_norm2(rt, n, x, s_); }
    void zaxpby(std::ptrdiff_t n, T a, const T *x, T b, const T *y, T *z) override {
        sfem_amd_blas_zaxpby(rt, n, a, x, b, y, z, s_);
    }
    void xypaz(std::ptrdiff_t n, const T *x, const T *y, T a, T *z) override {
        sfem_amd_blas_xypaz(rt, n, x, y, a, z, s_);
    }
    void *stream() const { return s_; }

private:
    void *s_;
};

template <typename T>
inline std::shared_ptr<BLAS<T>> make_device_blas(void *stream = nullptr) {
    return std::make_shared<DeviceBLAS<T>>(stream);
}

}  // namespace sfem
