// Quantum phase estimation (declaration as in the reference's
// include/lqcs/algorithms/qpe.hpp:7-23; implemented by the reference's
// src/algorithms/qpe.cpp over this simulator).
#pragma once

#include <cstdint>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct QPEResult {
    double phase;            // estimated phase in [0, 1)
    std::uint64_t measured;  // most likely counting-register value m, phase ~ m / 2^t
    std::size_t n_counting;
    QuantumCircuit circuit;
};

QPEResult phase_estimation(const Gate& u, const QuantumCircuit& eigenstate_prep, std::size_t n_counting);

}  // namespace lqcs::algorithms
