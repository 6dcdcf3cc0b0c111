// Expectation values, gradients and the two variational solvers (reference
// src/algorithms/vqe.cpp:11-133) and the QFT builder (src/algorithms/qft.cpp:7-27),
// on the engine.
#include <cmath>
#include <numbers>
#include <optional>
#include <stdexcept>

#include "lowering.hpp"
#include "lqcs/algorithms/qft.hpp"
#include "lqcs/algorithms/vqe.hpp"
#include "lqcs/backend/statevector_simulator.hpp"

namespace lqcs::algorithms {

using detail::check;

double expectation(const Statevector& sv, const Hamiltonian& hamiltonian) {
    std::vector<uint64_t> xs, zs;
    std::vector<uint32_t> ys;
    std::vector<double> cs;
    for (const PauliTerm& t : hamiltonian) {
        const detail::PauliString p = detail::parse_pauli(t.pauli, sv.num_qubits());
        xs.push_back(p.x);
        zs.push_back(p.zy);
        ys.push_back(p.n_y);
        cs.push_back(t.coeff);
    }
    double e = 0.0;
    check(lqcs_state_expectation_terms(sv.device(), xs.data(), zs.data(), ys.data(), cs.data(), cs.size(), &e));
    return e;
}

std::vector<double> parameter_shift_gradient(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz,
                                             std::span<const double> params) {
    const std::size_t n = ansatz.num_parameters();
    if (params.size() != n)
        throw std::invalid_argument("parameter_shift_gradient: params size must equal num_parameters");
    StatevectorSimulator sim;
    auto energy = [&](const std::vector<double>& th) {
        return expectation(sim.run_statevector(ansatz.bind(th)), hamiltonian);
    };
    constexpr double kHalfPi = std::numbers::pi / 2;
    std::vector<double> th(params.begin(), params.end()), grad(n);
    for (std::size_t k = 0; k < n; ++k) {
        const double saved = th[k];
        th[k] = saved + kHalfPi;
        const double ep = energy(th);
        th[k] = saved - kHalfPi;
        const double em = energy(th);
        th[k] = saved;
        grad[k] = 0.5 * (ep - em);
    }
    return grad;
}

namespace {

// Bound circuit as an op stream whose parameter gates are Pauli rotations.
struct Lowered {
    detail::ProgramBuilder prog;
    std::vector<uint64_t> param_op;
    std::vector<uint64_t> xs, zs;
    std::vector<uint32_t> ys;
    std::vector<double> cs;
};

void lower(const ParametricCircuit& ansatz, const Hamiltonian& h, std::span<const double> params, Lowered& L) {
    if (params.size() != ansatz.num_parameters())
        throw std::invalid_argument("adjoint_gradient: params size must equal num_parameters");
    const QuantumCircuit bound = ansatz.bind(params);
    std::vector<std::int64_t> param_at(bound.num_instructions(), -1);
    for (std::size_t p = 0; p < ansatz.num_parameters(); ++p)
        param_at[ansatz.parameter_instruction(p)] = static_cast<std::int64_t>(p);
    L.param_op.assign(ansatz.num_parameters(), 0);
    const auto insts = bound.instructions();
    for (std::size_t i = 0; i < insts.size(); ++i) {
        const Instruction& in = insts[i];
        if (in.gate.type == GateType::Reset)
            throw std::invalid_argument("run_statevector: Reset is non-deterministic; use run() instead");
        if (!in.gate.is_unitary()) continue;
        if (param_at[i] >= 0) {
            const detail::PauliString g = detail::generator_of(in);
            L.param_op[param_at[i]] = L.prog.add_pauli_rotation(g.x, g.zy, g.n_y, in.gate.params.at(0));
        } else {
            L.prog.add_instruction(in);
        }
    }
    for (const PauliTerm& t : h) {
        const detail::PauliString p = detail::parse_pauli(t.pauli, bound.num_qubits());
        L.xs.push_back(p.x);
        L.zs.push_back(p.zy);
        L.ys.push_back(p.n_y);
        L.cs.push_back(t.coeff);
    }
}

}  // namespace

std::vector<double> adjoint_gradient(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz,
                                     std::span<const double> params, double* energy) {
    Lowered L;
    lower(ansatz, hamiltonian, params, L);
    std::vector<double> grad(L.param_op.size());
    double e = 0.0;
    const lqcs_program prog = L.prog.program();
    check(lqcs_adjoint_gradient(static_cast<uint32_t>(ansatz.num_qubits()), &prog, L.param_op.data(),
                                L.param_op.size(), L.xs.data(), L.zs.data(), L.ys.data(), L.cs.data(), L.cs.size(),
                                &e, grad.data()));
    if (energy) *energy = e;
    return grad;
}

std::vector<GradientResult> adjoint_gradient_batched(std::span<const GradientJob> jobs) {
    std::vector<Lowered> L(jobs.size());
    std::vector<uint32_t> nq;
    std::vector<lqcs_program> progs;
    std::vector<uint64_t> param_op, param_off{0}, xs, zs, term_off{0};
    std::vector<uint32_t> ys;
    std::vector<double> cs;
    for (std::size_t c = 0; c < jobs.size(); ++c) {
        lower(*jobs[c].ansatz, *jobs[c].hamiltonian, jobs[c].params, L[c]);
        nq.push_back(static_cast<uint32_t>(jobs[c].ansatz->num_qubits()));
        param_op.insert(param_op.end(), L[c].param_op.begin(), L[c].param_op.end());
        param_off.push_back(param_op.size());
        xs.insert(xs.end(), L[c].xs.begin(), L[c].xs.end());
        zs.insert(zs.end(), L[c].zs.begin(), L[c].zs.end());
        ys.insert(ys.end(), L[c].ys.begin(), L[c].ys.end());
        cs.insert(cs.end(), L[c].cs.begin(), L[c].cs.end());
        term_off.push_back(cs.size());
    }
    for (auto& l : L) progs.push_back(l.prog.program());
    std::vector<double> energies(jobs.size()), grads(param_op.size());
    check(lqcs_adjoint_gradient_batched(static_cast<uint32_t>(jobs.size()), nq.data(), progs.data(),
                                        param_op.data(), param_off.data(), xs.data(), zs.data(), ys.data(),
                                        cs.data(), term_off.data(), energies.data(), grads.data()));
    std::vector<GradientResult> out(jobs.size());
    for (std::size_t c = 0; c < jobs.size(); ++c) {
        out[c].energy = energies[c];
        out[c].grad.assign(grads.begin() + param_off[c], grads.begin() + param_off[c + 1]);
    }
    return out;
}

AdjointPlan::AdjointPlan(std::span<const GradientJob> jobs) {
    if (jobs.empty()) throw std::invalid_argument("AdjointPlan: no jobs");
    std::vector<Lowered> L(jobs.size());
    std::vector<uint32_t> nq;
    std::vector<lqcs_program> progs;
    std::vector<uint64_t> param_op, xs, zs, term_off{0};
    std::vector<uint32_t> ys;
    std::vector<double> cs;
    param_off_.assign(1, 0);
    for (std::size_t c = 0; c < jobs.size(); ++c) {
        lower(*jobs[c].ansatz, *jobs[c].hamiltonian, jobs[c].params, L[c]);
        nq.push_back(static_cast<uint32_t>(jobs[c].ansatz->num_qubits()));
        param_op.insert(param_op.end(), L[c].param_op.begin(), L[c].param_op.end());
        param_off_.push_back(param_op.size());
        xs.insert(xs.end(), L[c].xs.begin(), L[c].xs.end());
        zs.insert(zs.end(), L[c].zs.begin(), L[c].zs.end());
        ys.insert(ys.end(), L[c].ys.begin(), L[c].ys.end());
        cs.insert(cs.end(), L[c].cs.begin(), L[c].cs.end());
        term_off.push_back(cs.size());
    }
    for (auto& l : L) progs.push_back(l.prog.program());
    check(lqcs_adjoint_plan_create(static_cast<uint32_t>(jobs.size()), nq.data(), progs.data(), param_op.data(),
                                   param_off_.data(), xs.data(), zs.data(), ys.data(), cs.data(), term_off.data(),
                                   &plan_));
}

AdjointPlan::AdjointPlan(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz)
    : AdjointPlan(std::span<const GradientJob>(
          std::vector<GradientJob>{{&ansatz, &hamiltonian, std::vector<double>(ansatz.num_parameters(), 0.0)}})) {}

AdjointPlan::~AdjointPlan() {
    if (plan_) (void)lqcs_adjoint_plan_destroy(plan_);
}

AdjointPlan::AdjointPlan(AdjointPlan&& o) noexcept : plan_(o.plan_), param_off_(std::move(o.param_off_)) {
    o.plan_ = nullptr;
}

AdjointPlan& AdjointPlan::operator=(AdjointPlan&& o) noexcept {
    if (this != &o) {
        if (plan_) (void)lqcs_adjoint_plan_destroy(plan_);
        plan_ = o.plan_;
        param_off_ = std::move(o.param_off_);
        o.plan_ = nullptr;
    }
    return *this;
}

std::vector<GradientResult> AdjointPlan::run(std::span<const std::vector<double>> params) {
    if (!plan_) throw std::logic_error("AdjointPlan: moved-from plan");
    const std::size_t n = num_jobs();
    if (params.size() != n) throw std::invalid_argument("AdjointPlan::run: one parameter vector per job");
    std::vector<double> theta;
    theta.reserve(param_off_.back());
    for (std::size_t c = 0; c < n; ++c) {
        if (params[c].size() != param_off_[c + 1] - param_off_[c])
            throw std::invalid_argument("AdjointPlan::run: params size must equal num_parameters");
        theta.insert(theta.end(), params[c].begin(), params[c].end());
    }
    std::vector<double> energies(n), grads(theta.size());
    check(lqcs_adjoint_plan_run(plan_, theta.data(), theta.size(), energies.data(), grads.data()));
    std::vector<GradientResult> out(n);
    for (std::size_t c = 0; c < n; ++c) {
        out[c].energy = energies[c];
        out[c].grad.assign(grads.begin() + param_off_[c], grads.begin() + param_off_[c + 1]);
    }
    return out;
}

std::vector<double> AdjointPlan::gradient(std::span<const double> params, double* energy) {
    if (!plan_) throw std::logic_error("AdjointPlan: moved-from plan");
    if (num_jobs() != 1) throw std::invalid_argument("AdjointPlan::gradient: the plan holds several jobs");
    if (params.size() != param_off_[1]) throw std::invalid_argument("AdjointPlan: params size must equal num_parameters");
    double e = 0.0;
    std::vector<double> g(params.size());
    check(lqcs_adjoint_plan_run(plan_, params.data(), params.size(), &e, g.data()));
    if (energy) *energy = e;
    return g;
}

// Rotosolve (reference src/algorithms/vqe.cpp:19-63). Along one coordinate
// E(t) = a + b cos(t - phi); with e0 = E(s), e+ = E(s + pi/2), e- = E(s - pi/2)
// the minimum sits at s - pi/2 - atan2(2 e0 - e+ - e-, e+ - e-), folded into
// (-pi, pi]. Each coordinate costs three state evolutions on the device.
VQEResult vqe(const Hamiltonian& hamiltonian, const Ansatz& ansatz, std::size_t n_params,
              const VQEOptions& options) {
    if (n_params == 0) throw std::invalid_argument("vqe: n_params must be >= 1");
    std::vector<double> theta = options.initial_parameters;
    if (theta.empty())
        theta.assign(n_params, 0.0);
    else if (theta.size() != n_params)
        throw std::invalid_argument("vqe: initial_parameters size must equal n_params");
    StatevectorSimulator sim;
    auto energy = [&](const std::vector<double>& th) {
        return expectation(sim.run_statevector(ansatz(std::span<const double>(th))), hamiltonian);
    };
    constexpr double kQuarter = std::numbers::pi / 2;
    VQEResult res{};
    double e = energy(theta);
    for (std::size_t sweep = 0; sweep < options.max_iterations; ++sweep) {
        const double start = e;
        for (double& t : theta) {
            const double s0 = t;
            t = s0 + kQuarter;
            const double ep = energy(theta);
            t = s0 - kQuarter;
            const double em = energy(theta);
            t = std::remainder(s0 - kQuarter - std::atan2(2.0 * e - ep - em, ep - em), 2.0 * std::numbers::pi);
            e = energy(theta);
        }
        res.history.push_back(e);
        res.iterations = sweep + 1;
        if (start - e < options.tol) break;
    }
    res.energy = e;
    res.parameters = std::move(theta);
    return res;
}

// Gradient descent (reference src/algorithms/vqe.cpp:96-133): theta -= lr * g,
// the energy after each step recorded, stopping once |g| < tol. The gradient
// is the adjoint one (one forward and one backward sweep instead of 2P
// evolutions); the adjoint call at the new point also returns the energy
// there, which is the step's history entry and the next step's gradient.
// The ansatz is lowered and its launch sequence captured once (AdjointPlan);
// an ansatz the plan refuses takes the per-call adjoint_gradient.
VQEResult vqe_gradient_descent(const Hamiltonian& hamiltonian, const ParametricCircuit& ansatz,
                               const GradientVQEOptions& options) {
    const std::size_t n = ansatz.num_parameters();
    if (n == 0) throw std::invalid_argument("vqe_gradient_descent: ansatz has no parameters");
    std::vector<double> theta = options.initial_parameters;
    if (theta.empty())
        theta.assign(n, 0.0);
    else if (theta.size() != n)
        throw std::invalid_argument("vqe_gradient_descent: initial_parameters size must equal num_parameters");
    VQEResult res{};
    std::optional<AdjointPlan> plan;
    try {
        plan.emplace(hamiltonian, ansatz);
    } catch (const std::invalid_argument&) {
    }
    double e = 0.0;
    auto grad_at = [&](const std::vector<double>& th) {
        return plan ? plan->gradient(th, &e) : adjoint_gradient(hamiltonian, ansatz, th, &e);
    };
    std::vector<double> g = grad_at(theta);
    for (std::size_t it = 0; it < options.max_iterations; ++it) {
        double norm2 = 0.0;
        for (std::size_t k = 0; k < n; ++k) {
            theta[k] -= options.learning_rate * g[k];
            norm2 += g[k] * g[k];
        }
        g = grad_at(theta);
        res.history.push_back(e);
        res.iterations = it + 1;
        if (std::sqrt(norm2) < options.tol) break;
    }
    res.energy = e;
    res.parameters = std::move(theta);
    return res;
}

QuantumCircuit qft(std::size_t n, bool do_swaps) {
    QuantumCircuit qc(n);
    for (std::size_t i = n; i-- > 0;) {
        qc.h(static_cast<qubit_t>(i));
        for (std::size_t j = i; j-- > 0;)
            qc.cp(std::numbers::pi / static_cast<double>(1ULL << (i - j)), static_cast<qubit_t>(j),
                  static_cast<qubit_t>(i));
    }
    if (do_swaps)
        for (std::size_t i = 0; i < n / 2; ++i) qc.swap(static_cast<qubit_t>(i), static_cast<qubit_t>(n - 1 - i));
    return qc;
}

QuantumCircuit qft_inverse(std::size_t n, bool do_swaps) { return qft(n, do_swaps).inverse(); }

}  // namespace lqcs::algorithms
