// sfem::Buffer<T>: owning or wrapping view of device memory (smesh Buffer, frontend/sfem_aliases.hpp:16-62).
#pragma once

#include "sfem_base.hpp"

namespace sfem {

template <typename T>
class Buffer {
public:
    Buffer(std::ptrdiff_t n, T *data, std::function<void(void *)> destroy)
        : n_(n), data_(data), destroy_(std::move(destroy)) {}
    ~Buffer() {
        if (destroy_ && data_) destroy_(data_);
    }
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;

    static std::shared_ptr<Buffer<T>> own(std::ptrdiff_t n, T *data, std::function<void(void *)> destroy) {
        return std::make_shared<Buffer<T>>(n, data, std::move(destroy));
    }
    static std::shared_ptr<Buffer<T>> wrap(std::ptrdiff_t n, T *data) {
        return std::make_shared<Buffer<T>>(n, data, nullptr);
    }
    T *data() { return data_; }
    const T *data() const { return data_; }
    std::ptrdiff_t size() const { return n_; }

private:
    std::ptrdiff_t n_;
    T *data_;
    std::function<void(void *)> destroy_;
};

template <typename T>
using SharedBuffer = std::shared_ptr<Buffer<T>>;

// Device buffer of n values (zero-initialised).
template <typename T>
inline SharedBuffer<T> create_device_buffer(std::ptrdiff_t n) {
    T *p = (T *)sfem_amd_malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
    if (!p) fatal("create_device_buffer");
    sfem_amd_memset(p, 0, sizeof(T) * (size_t)n, nullptr);
    return Buffer<T>::own(n, p, [](void *q) { sfem_amd_free(q); });
}

}  // namespace sfem
