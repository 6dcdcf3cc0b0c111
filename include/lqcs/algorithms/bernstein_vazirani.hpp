// Bernstein-Vazirani (declaration as in the reference's
// include/lqcs/algorithms/bernstein_vazirani.hpp:8-20; implemented by the
// reference's src/algorithms/bernstein_vazirani.cpp over this simulator).
#pragma once

#include <cstdint>
#include <functional>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::algorithms {

struct BernsteinVaziraniResult {
    std::uint64_t secret;  // the hidden string s recovered from f(x) = s.x mod 2
    QuantumCircuit circuit;
};

BernsteinVaziraniResult bernstein_vazirani(std::size_t n, const std::function<bool(std::uint64_t)>& f);

}  // namespace lqcs::algorithms
