// ref_driver.cpp -- C entry points over the REFERENCE implementation itself,
// compiled from its own sources by oracle/build_ref.sh into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it pins the C restatement in lqcs_oracle.c to the
// reference and times the reference's OpenMP path as the CPU baseline. It uses
// the reference's public API (QuantumCircuit::append, StatevectorSimulator,
// ParametricCircuit, algorithms::expectation/parameter_shift_gradient,
// DensityMatrixSimulator) and its private kernel header for kernel-level checks.
// Circuits arrive in the oracle's flat format (lqcs_oracle.h).
#include <omp.h>

#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include <lqcs/lqcs.hpp>

#include "backend/kernels/kernels.hpp"
#include "lqcs_oracle.h"

using namespace lqcs;

namespace {

thread_local std::string g_err;

Instruction to_instruction(const orc_inst& in) {
    std::vector<double> params(in.params, in.params + in.n_params);
    std::vector<complex_t> mat;
    for (uint64_t i = 0; i < in.mat_len; ++i) mat.emplace_back(in.mat[i].re, in.mat[i].im);
    std::vector<std::size_t> perm(in.perm, in.perm + in.perm_len);
    std::string pauli = in.pauli ? std::string(in.pauli, in.pauli_len) : std::string();
    Gate g(static_cast<GateType>(in.type), std::move(params), in.n_controls, std::move(mat),
           std::move(perm), std::move(pauli));
    std::vector<qubit_t> qubits(in.qubits, in.qubits + in.n_qubits);
    std::vector<clbit_t> clbits;
    if (in.type == static_cast<int32_t>(GateType::Measure)) clbits.push_back(in.clbit);
    return Instruction{std::move(g), std::move(qubits), std::move(clbits)};
}

QuantumCircuit to_circuit(const orc_circuit* c) {
    QuantumCircuit qc(c->n_qubits, c->n_clbits);
    for (uint64_t i = 0; i < c->n_inst; ++i) qc.append(to_instruction(c->inst[i]));
    return qc;
}

NoiseModel to_noise(const orc_noise* n) {
    NoiseModel nm;
    if (!n) return nm;
    for (uint32_t k = 0; k < n->n_channels; ++k) {
        std::vector<std::array<complex_t, 4>> ops;
        for (uint32_t j = 0; j < n->channels[k].n_ops; ++j) {
            std::array<complex_t, 4> a;
            for (int e = 0; e < 4; ++e) {
                const orc_c z = n->channels[k].ops[4 * j + e];
                a[e] = complex_t(z.re, z.im);
            }
            ops.push_back(a);
        }
        nm.add_all_qubit_channel(KrausChannel(std::move(ops)));
    }
    return nm;
}

algorithms::Hamiltonian to_hamiltonian(const orc_term* h, uint64_t nt) {
    algorithms::Hamiltonian H;
    for (uint64_t t = 0; t < nt; ++t) H.push_back({h[t].coeff, std::string(h[t].pauli)});
    return H;
}

// ParametricCircuit is built only through its own builder methods.
ParametricCircuit to_parametric(const orc_circuit* c, const uint64_t* param_inst, uint64_t np) {
    ParametricCircuit pc(c->n_qubits, c->n_clbits);
    std::vector<int> is_param(c->n_inst, 0);
    for (uint64_t p = 0; p < np; ++p) is_param[param_inst[p]] = 1;
    for (uint64_t i = 0; i < c->n_inst; ++i) {
        const orc_inst& g = c->inst[i];
        const auto t = static_cast<GateType>(g.type);
        const qubit_t* q = g.qubits;
        if (is_param[i]) {
            if (t == GateType::RX) pc.rx(q[0]);
            else if (t == GateType::RY) pc.ry(q[0]);
            else if (t == GateType::RZ) pc.rz(q[0]);
            else if (t == GateType::PauliRotation)
                pc.rp(std::string_view(g.pauli, g.pauli_len),
                      std::span<const qubit_t>(q, g.n_qubits));
            else throw std::invalid_argument("ref_driver: unsupported parametric gate");
            continue;
        }
        if (g.n_controls == 0) {
            switch (t) {
                case GateType::X: pc.x(q[0]); break;
                case GateType::Y: pc.y(q[0]); break;
                case GateType::Z: pc.z(q[0]); break;
                case GateType::H: pc.h(q[0]); break;
                case GateType::S: pc.s(q[0]); break;
                case GateType::Sdg: pc.sdg(q[0]); break;
                case GateType::T: pc.t(q[0]); break;
                case GateType::SX: pc.sx(q[0]); break;
                case GateType::SWAP: pc.swap(q[0], q[1]); break;
                case GateType::Barrier: pc.barrier(); break;
                default: throw std::invalid_argument("ref_driver: unsupported fixed gate");
            }
        } else if (g.n_controls == 1 && t == GateType::X) {
            pc.cx(q[0], q[1]);
        } else if (g.n_controls == 1 && t == GateType::Y) {
            pc.cy(q[0], q[1]);
        } else if (g.n_controls == 1 && t == GateType::Z) {
            pc.cz(q[0], q[1]);
        } else if (g.n_controls == 2 && t == GateType::X) {
            pc.ccx(q[0], q[1], q[2]);
        } else {
            throw std::invalid_argument("ref_driver: unsupported controlled gate");
        }
    }
    return pc;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::out_of_range& e) {
        g_err = e.what();
        return 2;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return 3;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 9;
    }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void ref_set_threads(int n) { omp_set_num_threads(n); }

int ref_run_statevector(const orc_circuit* c, int fuse, orc_c* out) {
    return guarded([&] {
        StatevectorSimulator::Options o;
        o.fuse_gates = fuse != 0;
        const Statevector sv = StatevectorSimulator(o).run_statevector(to_circuit(c));
        std::memcpy(out, sv.data(), sizeof(orc_c) * sv.size());
    });
}

// Counts of StatevectorSimulator::run, written as "key=count\n" lines.
int ref_run_counts(const orc_circuit* c, const orc_noise* noise, int fuse, uint64_t seed,
                   uint64_t shots, char* buf, uint64_t buf_len) {
    return guarded([&] {
        StatevectorSimulator::Options o;
        o.seed = seed;
        o.fuse_gates = fuse != 0;
        o.noise = to_noise(noise);
        const Result r = StatevectorSimulator(o).run(to_circuit(c), shots);
        std::string s;
        for (const auto& [k, v] : r.counts()) s += k + "=" + std::to_string(v) + "\n";
        if (s.size() + 1 > buf_len) throw std::length_error("ref_run_counts: buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int ref_expectation_value(const orc_c* state, uint32_t nq, const char* pauli, double* out) {
    return guarded([&] {
        Statevector sv(nq);
        std::memcpy(sv.data(), state, sizeof(orc_c) * sv.size());
        *out = sv.expectation_value(pauli);
    });
}

int ref_expectation(const orc_c* state, uint32_t nq, const orc_term* h, uint64_t nt,
                    double* out) {
    return guarded([&] {
        Statevector sv(nq);
        std::memcpy(sv.data(), state, sizeof(orc_c) * sv.size());
        *out = algorithms::expectation(sv, to_hamiltonian(h, nt));
    });
}

int ref_parameter_shift_gradient(const orc_circuit* tmpl, const uint64_t* param_inst,
                                 uint64_t np, const double* theta, const orc_term* h,
                                 uint64_t nt, double* grad) {
    return guarded([&] {
        const ParametricCircuit pc = to_parametric(tmpl, param_inst, np);
        const auto g = algorithms::parameter_shift_gradient(
            to_hamiltonian(h, nt), pc, std::span<const double>(theta, np));
        std::memcpy(grad, g.data(), sizeof(double) * np);
    });
}

int ref_bind_statevector(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t np,
                         const double* theta, orc_c* out) {
    return guarded([&] {
        const ParametricCircuit pc = to_parametric(tmpl, param_inst, np);
        const Statevector sv =
            StatevectorSimulator().run_statevector(pc.bind(std::span<const double>(theta, np)));
        std::memcpy(out, sv.data(), sizeof(orc_c) * sv.size());
    });
}

int ref_density_probabilities(const orc_circuit* c, const orc_noise* noise, double* out) {
    return guarded([&] {
        DensityMatrixSimulator::Options o;
        o.noise = to_noise(noise);
        const DensityMatrix rho = DensityMatrixSimulator(o).run_density_matrix(to_circuit(c));
        const auto p = rho.probabilities();
        std::memcpy(out, p.data(), sizeof(double) * p.size());
    });
}

int ref_kraus_probability(const orc_c* state, uint64_t n, uint32_t t, const orc_c* k,
                          double* out) {
    return guarded([&] {
        std::array<complex_t, 4> kk;
        for (int e = 0; e < 4; ++e) kk[e] = complex_t(k[e].re, k[e].im);
        *out = kernels::kraus_probability(reinterpret_cast<const complex_t*>(state), n, t,
                                          kk.data());
    });
}

// Gate::base_matrix of one instruction (complex dim*dim, row major).
int ref_base_matrix(const orc_inst* g, orc_c* out, uint64_t* dim) {
    return guarded([&] {
        const Instruction inst = to_instruction(*g);
        const auto m = inst.gate.base_matrix();
        std::memcpy(out, m.data(), sizeof(orc_c) * m.size());
        *dim = std::size_t{1} << inst.gate.base_qubits();
    });
}

namespace {
void put_text(const std::string& s, char* buf, uint64_t buf_len) {
    if (s.size() + 1 > buf_len) throw std::length_error("ref_driver: text buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
}
}  // namespace

// io::to_qasm of a circuit.
int ref_to_qasm(const orc_circuit* c, char* buf, uint64_t buf_len) {
    return guarded([&] { put_text(io::to_qasm(to_circuit(c)), buf, buf_len); });
}

// io::from_qasm of `text`, written back out by io::to_qasm (the exporter is
// pinned on its own, so the text is a faithful record of what was parsed).
int ref_qasm_reparse(const char* text, char* buf, uint64_t buf_len, uint64_t* n_inst) {
    return guarded([&] {
        const QuantumCircuit qc = io::from_qasm(text);
        *n_inst = qc.num_instructions();
        put_text(io::to_qasm(qc), buf, buf_len);
    });
}

// RedundantGateElimination::run, as QASM text.
int ref_eliminate_qasm(const orc_circuit* c, char* buf, uint64_t buf_len, uint64_t* n_inst) {
    return guarded([&] {
        const QuantumCircuit out = RedundantGateElimination().run(to_circuit(c));
        *n_inst = out.num_instructions();
        put_text(io::to_qasm(out), buf, buf_len);
    });
}

// algorithms::qft / qft_inverse, as QASM text.
int ref_qft_qasm(uint32_t n, int do_swaps, int inverse, char* buf, uint64_t buf_len) {
    return guarded([&] {
        const QuantumCircuit qc = inverse ? algorithms::qft_inverse(n, do_swaps != 0)
                                          : algorithms::qft(n, do_swaps != 0);
        put_text(io::to_qasm(qc), buf, buf_len);
    });
}

// algorithms::vqe (Rotosolve) with the ansatz theta -> template.bind(theta).
// history must hold max_iterations doubles.
int ref_vqe_rotosolve(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t np,
                      const orc_term* h, uint64_t nt, uint64_t max_iterations, double tol,
                      const double* init, double* energy, double* params, uint64_t* iterations,
                      double* history) {
    return guarded([&] {
        const ParametricCircuit pc = to_parametric(tmpl, param_inst, np);
        algorithms::VQEOptions o;
        o.max_iterations = max_iterations;
        o.tol = tol;
        if (init) o.initial_parameters.assign(init, init + np);
        const algorithms::Ansatz ansatz = [&pc](std::span<const double> th) { return pc.bind(th); };
        const auto r = algorithms::vqe(to_hamiltonian(h, nt), ansatz, np, o);
        *energy = r.energy;
        std::memcpy(params, r.parameters.data(), sizeof(double) * np);
        *iterations = r.iterations;
        std::memcpy(history, r.history.data(), sizeof(double) * r.history.size());
    });
}

// algorithms::vqe_gradient_descent. history must hold max_iterations doubles.
int ref_vqe_gradient_descent(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t np,
                             const orc_term* h, uint64_t nt, uint64_t max_iterations,
                             double learning_rate, double tol, const double* init, double* energy,
                             double* params, uint64_t* iterations, double* history) {
    return guarded([&] {
        const ParametricCircuit pc = to_parametric(tmpl, param_inst, np);
        algorithms::GradientVQEOptions o;
        o.max_iterations = max_iterations;
        o.learning_rate = learning_rate;
        o.tol = tol;
        if (init) o.initial_parameters.assign(init, init + np);
        const auto r = algorithms::vqe_gradient_descent(to_hamiltonian(h, nt), pc, o);
        *energy = r.energy;
        std::memcpy(params, r.parameters.data(), sizeof(double) * np);
        *iterations = r.iterations;
        std::memcpy(history, r.history.data(), sizeof(double) * r.history.size());
    });
}

// Draws of lqcs::Rng (mt19937_64 + uniform_real_distribution) for a seed.
void ref_uniform_stream(uint64_t seed, uint64_t n, double* out) {
    Rng r(seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = r.uniform();
}

}  // extern "C"
