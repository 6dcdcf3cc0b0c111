// Modular-exponentiation circuit builder (declaration as in the reference's
// include/lqcs/algorithms/arithmetic.hpp:8-20). Implemented by the reference's
// own src/algorithms/arithmetic.cpp, which compiles unchanged against these
// headers; its controlled permutations run on the engine's permutation path.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

// Counting register of n_counting qubits controlling |y> -> |a^(2^j) y mod N>
// on a work register of n_work qubits.
QuantumCircuit modular_exponentiation(std::uint64_t a, std::uint64_t N, std::size_t n_counting,
                                      std::size_t n_work);

}  // namespace lqcs::algorithms
