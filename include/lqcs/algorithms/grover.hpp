// Grover search (declaration as in the reference's
// include/lqcs/algorithms/grover.hpp:9-23; implemented by the reference's
// src/algorithms/grover.cpp over this simulator).
#pragma once

#include <cstdint>
#include <optional>
#include <span>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct GroverResult {
    std::uint64_t best_state;    // most probable basis state of the final state
    double success_probability;  // probability of measuring any marked state
    std::size_t iterations;      // Grover iterations applied
    QuantumCircuit circuit;      // ends in measure_all, ready for run()
};

// Without `iterations`, the optimal count for |marked| solutions is used.
GroverResult grover(std::size_t n, std::span<const std::uint64_t> marked,
                    std::optional<std::size_t> iterations = std::nullopt);

}  // namespace lqcs::algorithms
