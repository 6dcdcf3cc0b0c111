// Quantum Fourier transform builder (reference src/algorithms/qft.cpp:7-27).
#pragma once

#include <cstddef>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

QuantumCircuit qft(std::size_t n, bool do_swaps = true);
QuantumCircuit qft_inverse(std::size_t n, bool do_swaps = true);

}  // namespace lqcs::algorithms
