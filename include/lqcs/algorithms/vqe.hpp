// Expectations, gradients and variational solvers of Pauli Hamiltonians
// (reference include/lqcs/algorithms/vqe.hpp:12-64), plus the adjoint gradient.
#pragma once

#include <functional>
#include <span>
#include <string>
#include <vector>

#include "../backend/statevector.hpp"
#include "../circuit/parametric_circuit.hpp"
#include "../circuit/quantum_circuit.hpp"

struct lqcs_adjoint_plan;  // the engine's captured plan (lqcs_device.h)

namespace lqcs::algorithms {

struct PauliTerm {
    double coeff;
    std::string pauli;  // leftmost char = highest qubit
};
using Hamiltonian = std::vector<PauliTerm>;

// <psi|H|psi>
double expectation(const Statevector& sv, const Hamiltonian& hamiltonian);

// Parameter vector -> circuit preparing |psi(theta)>.
using Ansatz = std::function<QuantumCircuit(std::span<const double>)>;

struct VQEOptions {
    std::size_t max_iterations = 100;        // sweeps over all parameters
    double tol = 1e-9;                       // stop when a sweep lowers E by less
    std::vector<double> initial_parameters;  // empty: all zero
};

struct VQEResult {
    double energy;                   // final (variational upper bound) energy
    std::vector<double> parameters;  // parameters reaching it
    std::size_t iterations;          // sweeps or steps taken
    std::vector<double> history;     // energy after each sweep or step
};

// Rotosolve: every parameter is a rotation angle, so E is a sinusoid in each
// coordinate and three energies fix its minimum.
VQEResult vqe(const Hamiltonian& hamiltonian, const Ansatz& ansatz, std::size_t n_params,
              const VQEOptions& options = {});

// dE/dtheta_k = [E(theta + pi/2 e_k) - E(theta - pi/2 e_k)] / 2, 2P evaluations.
std::vector<double> parameter_shift_gradient(const Hamiltonian& hamiltonian,
                                             const ParametricCircuit& ansatz,
                                             std::span<const double> params);

struct GradientVQEOptions {
    std::size_t max_iterations = 200;
    double learning_rate = 0.1;
    double tol = 1e-9;                       // stop when the gradient norm is below
    std::vector<double> initial_parameters;  // empty: all zero
};

// Gradient descent on the exact gradient (computed here by the adjoint
// method, equal to the parameter-shift gradient).
VQEResult vqe_gradient_descent(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz,
                               const GradientVQEOptions& options = {});

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

// Engine extension: the jobs of adjoint_gradient_batched with their circuits
// and Hamiltonians fixed and their parameter values free, for optimiser
// loops. Construction lowers every job once and captures the device launch
// sequence in a graph (the jobs' params only fix the sizes); run() evaluates
// every job at new values by replaying it, with the same results as
// adjoint_gradient_batched. The ansatz and Hamiltonian objects are not used
// after construction. Ansatze with multi-controlled or permutation gates are
// refused (std::invalid_argument).
class AdjointPlan {
public:
    explicit AdjointPlan(std::span<const GradientJob> jobs);
    AdjointPlan(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz);
    ~AdjointPlan();
    AdjointPlan(AdjointPlan&& o) noexcept;
    AdjointPlan& operator=(AdjointPlan&& o) noexcept;
    AdjointPlan(const AdjointPlan&) = delete;
    AdjointPlan& operator=(const AdjointPlan&) = delete;

    std::size_t num_jobs() const { return param_off_.empty() ? 0 : param_off_.size() - 1; }
    // params[c] holds job c's values.
    std::vector<GradientResult> run(std::span<const std::vector<double>> params);
    // One-job plans: the gradient at `params`, the energy into `energy`.
    std::vector<double> gradient(std::span<const double> params, double* energy = nullptr);

private:
    ::lqcs_adjoint_plan* plan_ = nullptr;
    std::vector<std::uint64_t> param_off_;
};

}  // namespace lqcs::algorithms
