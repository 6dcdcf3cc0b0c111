// Oracle circuit builders (declarations as in the reference's
// include/lqcs/algorithms/oracles.hpp:9-20; implemented by the reference's
// src/algorithms/oracles.cpp, as controlled permutations and phases).
#pragma once

#include <cstdint>
#include <functional>
#include <span>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

// |x>|y> -> |x>|y xor f(x)> on n_in + n_out qubits.
QuantumCircuit xor_oracle(std::size_t n_in, std::size_t n_out,
                          const std::function<std::uint64_t(std::uint64_t)>& f);
// |x> -> -|x> for x in marked.
QuantumCircuit phase_oracle(std::size_t n, std::span<const std::uint64_t> marked);

}  // namespace lqcs::algorithms
