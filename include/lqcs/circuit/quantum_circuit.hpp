// Circuit IR with build-time validation; the fluent builder of the
// reference's include/lqcs/circuit/quantum_circuit.hpp:17-119.
#pragma once

#include <cstddef>
#include <initializer_list>
#include <span>
#include <string>
#include <string_view>
#include <vector>

#include "instruction.hpp"

namespace lqcs {

class QuantumCircuit {
public:
    explicit QuantumCircuit(std::size_t num_qubits, std::size_t num_clbits = 0);

    QuantumCircuit& i(qubit_t q);
    QuantumCircuit& x(qubit_t q);
    QuantumCircuit& y(qubit_t q);
    QuantumCircuit& z(qubit_t q);
    QuantumCircuit& h(qubit_t q);
    QuantumCircuit& s(qubit_t q);
    QuantumCircuit& sdg(qubit_t q);
    QuantumCircuit& t(qubit_t q);
    QuantumCircuit& tdg(qubit_t q);
    QuantumCircuit& sx(qubit_t q);

    QuantumCircuit& rx(double theta, qubit_t q);
    QuantumCircuit& ry(double theta, qubit_t q);
    QuantumCircuit& rz(double theta, qubit_t q);
    QuantumCircuit& p(double lambda, qubit_t q);
    QuantumCircuit& u(double theta, double phi, double lambda, qubit_t q);

    QuantumCircuit& cx(qubit_t control, qubit_t target);
    QuantumCircuit& cy(qubit_t control, qubit_t target);
    QuantumCircuit& cz(qubit_t control, qubit_t target);
    QuantumCircuit& ch(qubit_t control, qubit_t target);
    QuantumCircuit& cp(double lambda, qubit_t control, qubit_t target);
    QuantumCircuit& crx(double theta, qubit_t control, qubit_t target);
    QuantumCircuit& cry(double theta, qubit_t control, qubit_t target);
    QuantumCircuit& crz(double theta, qubit_t control, qubit_t target);
    QuantumCircuit& ccx(qubit_t c1, qubit_t c2, qubit_t target);
    QuantumCircuit& cswap(qubit_t control, qubit_t a, qubit_t b);
    QuantumCircuit& mcx(std::span<const qubit_t> controls, qubit_t target);
    QuantumCircuit& mcx(std::initializer_list<qubit_t> controls, qubit_t target);
    QuantumCircuit& mcp(double lambda, std::span<const qubit_t> controls, qubit_t target);
    QuantumCircuit& mcp(double lambda, std::initializer_list<qubit_t> controls, qubit_t target);

    QuantumCircuit& swap(qubit_t a, qubit_t b);
    QuantumCircuit& iswap(qubit_t a, qubit_t b);
    QuantumCircuit& rxx(double theta, qubit_t a, qubit_t b);
    QuantumCircuit& ryy(double theta, qubit_t a, qubit_t b);
    QuantumCircuit& rzz(double theta, qubit_t a, qubit_t b);

    QuantumCircuit& rp(double theta, std::string_view pauli, std::span<const qubit_t> qubits);
    QuantumCircuit& rp(double theta, std::string_view pauli, std::initializer_list<qubit_t> qubits);

    QuantumCircuit& unitary(std::span<const complex_t> matrix, std::span<const qubit_t> qubits);
    QuantumCircuit& unitary(std::span<const complex_t> matrix, std::initializer_list<qubit_t> qubits);
    QuantumCircuit& permutation(std::span<const std::size_t> table, std::span<const qubit_t> qubits);
    QuantumCircuit& permutation(std::span<const std::size_t> table, std::initializer_list<qubit_t> qubits);

    QuantumCircuit& measure(qubit_t q, clbit_t c);
    QuantumCircuit& measure_all();
    QuantumCircuit& reset(qubit_t q);
    QuantumCircuit& barrier();

    QuantumCircuit& append(Instruction inst);

    QuantumCircuit& compose(const QuantumCircuit& other, std::span<const qubit_t> qubit_map = {});
    QuantumCircuit inverse() const;
    QuantumCircuit power(int n) const;

    std::size_t num_qubits() const { return num_qubits_; }
    std::size_t num_clbits() const { return num_clbits_; }
    std::size_t num_instructions() const { return instructions_.size(); }
    std::span<const Instruction> instructions() const { return instructions_; }
    std::size_t depth() const;

    std::string draw() const;

private:
    QuantumCircuit& add_gate(Gate gate, std::vector<qubit_t> qubits);
    void check_qubit(qubit_t q) const;
    void check_clbit(clbit_t c) const;

    std::size_t num_qubits_;
    std::size_t num_clbits_;
    std::vector<Instruction> instructions_;
};

}  // namespace lqcs
