// Deutsch-Jozsa (declaration as in the reference's
// include/lqcs/algorithms/deutsch_jozsa.hpp:8-21; implemented by the
// reference's src/algorithms/deutsch_jozsa.cpp over this simulator).
#pragma once

#include <cstdint>
#include <functional>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct DeutschJozsaResult {
    bool is_constant;   // constant or balanced f
    double p_all_zero;  // probability of reading all zeros on the input register (1 or 0)
    QuantumCircuit circuit;
};

DeutschJozsaResult deutsch_jozsa(std::size_t n, const std::function<bool(std::uint64_t)>& f);

}  // namespace lqcs::algorithms
