// Seedable random source (the reference's include/lqcs/core/random.hpp:10-23).
// The draw is stated explicitly rather than delegated to the standard
// library's uniform_real_distribution: one mt19937_64 word w maps to
// double(w) / 2^64, clamped below 1 -- the value libstdc++ produces -- so the
// draw stream is the same on every host.
#pragma once

#include <cmath>
#include <cstdint>
#include <optional>
#include <random>

namespace lqcs {

class Rng {
public:
    explicit Rng(std::optional<std::uint64_t> seed = std::nullopt)
        : engine_(seed ? *seed : std::random_device{}()) {}

    double uniform() {
        const double v = static_cast<double>(engine_()) * 0x1p-64;
        return v >= 1.0 ? std::nextafter(1.0, 0.0) : v;
    }

    std::mt19937_64& engine() { return engine_; }

private:
    std::mt19937_64 engine_;
};

}  // namespace lqcs
