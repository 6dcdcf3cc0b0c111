// Shor factoring (declarations as in the reference's
// include/lqcs/algorithms/shor.hpp:10-35; implemented by the reference's
// src/algorithms/shor.cpp over this simulator).
#pragma once

#include <cstdint>
#include <map>
#include <optional>
#include <string>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct ShorOptions {
    std::size_t shots = 64;             // samples per attempt
    std::size_t max_attempts = 10;      // random bases a to try
    std::optional<std::uint64_t> seed;  // fixes the whole run when set
};

struct ShorResult {
    bool success = false;
    std::uint64_t N = 0;
    std::uint64_t factor1 = 0, factor2 = 0;
    std::uint64_t a = 0;          // base used (or the classical shortcut's)
    std::uint64_t period = 0;     // period r found (0 for a classical shortcut)
    std::size_t attempts = 0;     // attempts made
    std::map<std::string, std::size_t> counts;  // counting-register histogram
    std::optional<QuantumCircuit> circuit;      // circuit of the successful attempt
};

// Classical shortcuts (even, prime, prime power, lucky gcd), then quantum
// period finding.
ShorResult shor(std::uint64_t N, const ShorOptions& options = {});

}  // namespace lqcs::algorithms
