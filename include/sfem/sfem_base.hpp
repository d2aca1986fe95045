// Base definitions of the sfem:: drop-in API (base/sfem_base.hpp, frontend/sfem_aliases.hpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>

#include "sfem_amd.h"

namespace sfem {

using real_t = double;  // cmake/SFEMNumbers.cmake:5-8 defaults
using geom_t = float;
using idx_t = int32_t;

// The library computes on the GPU: operators and vectors of an object live in one space.
enum ExecutionSpace { EXECUTION_SPACE_HOST = 0, EXECUTION_SPACE_DEVICE = 1, EXECUTION_SPACE_INVALID = -1 };

template <typename T>
struct real_type_of;
template <>
struct real_type_of<double> {
    static constexpr int value = SFEM_AMD_FLOAT64;
};
template <>
struct real_type_of<float> {
    static constexpr int value = SFEM_AMD_FLOAT32;
};

// Fatal errors abort like SFEM_ERROR (base/sfem_base.hpp:58-65).
[[noreturn]] inline void fatal(const char *what) {
    std::fprintf(stderr, "[sfem_amd] %s: %s\n", what, sfem_amd_last_error());
    std::abort();
}

}  // namespace sfem
