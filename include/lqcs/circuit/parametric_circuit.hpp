// Parametric circuit: a template with parameter slots on RX/RY/RZ/
// PauliRotation gates (reference include/lqcs/circuit/parametric_circuit.hpp:20-63).
#pragma once

#include <cstddef>
#include <initializer_list>
#include <span>
#include <string_view>
#include <vector>

#include "quantum_circuit.hpp"

namespace lqcs {

class ParametricCircuit {
public:
    explicit ParametricCircuit(std::size_t num_qubits, std::size_t num_clbits = 0);

    ParametricCircuit& x(qubit_t q);
    ParametricCircuit& y(qubit_t q);
    ParametricCircuit& z(qubit_t q);
    ParametricCircuit& h(qubit_t q);
    ParametricCircuit& s(qubit_t q);
    ParametricCircuit& sdg(qubit_t q);
    ParametricCircuit& t(qubit_t q);
    ParametricCircuit& sx(qubit_t q);
    ParametricCircuit& cx(qubit_t control, qubit_t target);
    ParametricCircuit& cy(qubit_t control, qubit_t target);
    ParametricCircuit& cz(qubit_t control, qubit_t target);
    ParametricCircuit& swap(qubit_t a, qubit_t b);
    ParametricCircuit& ccx(qubit_t c1, qubit_t c2, qubit_t target);
    ParametricCircuit& barrier();

    std::size_t rx(qubit_t q);
    std::size_t ry(qubit_t q);
    std::size_t rz(qubit_t q);
    std::size_t rp(std::string_view pauli, std::span<const qubit_t> qubits);
    std::size_t rp(std::string_view pauli, std::initializer_list<qubit_t> qubits);

    QuantumCircuit bind(std::span<const double> values) const;
    QuantumCircuit bind(std::initializer_list<double> values) const {
        return bind(std::span<const double>(values.begin(), values.size()));
    }

    std::size_t num_parameters() const { return param_inst_.size(); }
    std::size_t num_qubits() const { return template_.num_qubits(); }
    std::size_t num_clbits() const { return template_.num_clbits(); }

    // Engine extension: the template and the instruction each parameter sets.
    const QuantumCircuit& circuit_template() const { return template_; }
    std::size_t parameter_instruction(std::size_t p) const { return param_inst_.at(p); }

private:
    std::size_t add_param_gate(Gate gate, std::vector<qubit_t> qubits);

    QuantumCircuit template_;
    std::vector<std::size_t> param_inst_;
};

}  // namespace lqcs
