// Simon's algorithm (declaration as in the reference's
// include/lqcs/algorithms/simon.hpp:9-24; implemented by the reference's
// src/algorithms/simon.cpp over this simulator).
#pragma once

#include <cstdint>
#include <functional>
#include <optional>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct SimonResult {
    std::uint64_t secret;         // period s with f(x) = f(x xor s)
    std::size_t quantum_samples;  // quantum samples used
    QuantumCircuit circuit;
};

SimonResult simon(std::size_t n, const std::function<std::uint64_t(std::uint64_t)>& f,
                  std::optional<std::uint64_t> seed = std::nullopt);

}  // namespace lqcs::algorithms
