// OpenQASM 2.0 export and import (reference include/lqcs/io/qasm.hpp:9-21).
//
// to_qasm writes the qelib1.inc gate set: every single-qubit gate (P as u1,
// U as u3), cx/cy/cz/ch/cu1/crx/cry/crz, swap/ccx/cswap, measure, reset and
// barrier, with parameters printed as %.17g so that they read back exactly.
// Gates outside that set (iSWAP, RXX/RYY/RZZ, Pauli rotations, Unitary,
// Permutation, three or more controls) throw std::invalid_argument.
//
// from_qasm reads that output and the common subset other tools write: one
// qreg, at most one creg (declared before any gate), comments, include lines,
// u2, whole-register measure, and parameter expressions over numbers and pi
// with + - * / and parentheses. Errors throw std::invalid_argument.
#pragma once

#include <string>
#include <string_view>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs::io {

std::string to_qasm(const QuantumCircuit& circuit);
QuantumCircuit from_qasm(std::string_view source);

}  // namespace lqcs::io
