// Index bit helpers (the reference's include/lqcs/core/bit_utils.hpp:9-16).
#pragma once

#include <cstddef>

namespace lqcs::bits {

// Inserts a 0 at bit `pos`: bits below stay, bits at and above move up one.
constexpr std::size_t insert_zero_bit(std::size_t value, unsigned pos) {
    const std::size_t below = (std::size_t{1} << pos) - 1;
    return ((value & ~below) << 1) | (value & below);
}

constexpr bool test_bit(std::size_t value, unsigned pos) { return ((value >> pos) & 1u) != 0; }

}  // namespace lqcs::bits
