// Expectations and gradients of Pauli Hamiltonians (reference
// include/lqcs/algorithms/vqe.hpp:15-52), plus the adjoint gradient.
#pragma once

#include <span>
#include <string>
#include <vector>

#include "../backend/statevector.hpp"
#include "../circuit/parametric_circuit.hpp"

namespace lqcs::algorithms {

struct PauliTerm {
    double coeff;
    std::string pauli;  // leftmost char = highest qubit
};
using Hamiltonian = std::vector<PauliTerm>;

// <psi|H|psi>
double expectation(const Statevector& sv, const Hamiltonian& hamiltonian);

// dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2, 2P evaluations.
std::vector<double> parameter_shift_gradient(const Hamiltonian& hamiltonian,
                                             const ParametricCircuit& ansatz,
                                             std::span<const double> params);

// Engine extension: the same gradient from one forward and one backward
// sweep (adjoint method). `energy`, when given, receives E(theta).
std::vector<double> adjoint_gradient(const Hamiltonian& hamiltonian,
                                     const ParametricCircuit& ansatz,
                                     std::span<const double> params, double* energy = nullptr);

// Many ansatz/Hamiltonian/parameter sets in shared device launches.
struct GradientJob {
    const ParametricCircuit* ansatz;
    const Hamiltonian* hamiltonian;
    std::vector<double> params;
};
struct GradientResult {
    double energy;
    std::vector<double> grad;
};
std::vector<GradientResult> adjoint_gradient_batched(std::span<const GradientJob> jobs);

}  // namespace lqcs::algorithms
