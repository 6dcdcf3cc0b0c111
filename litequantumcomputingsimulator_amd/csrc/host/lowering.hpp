// Host-side lowering of circuit instructions to the device op stream, and the
// mapping of C-ABI status codes back to the exception types callers expect.
#pragma once

#include <cstdint>
#include <string>
#include <string_view>
#include <vector>

#include "lqcs/circuit/instruction.hpp"
#include "lqcs_device.h"

namespace lqcs::detail {

// Throws the exception type of `status` (invalid_argument, out_of_range,
// logic_error, bad_alloc-like runtime_error, ...) with the engine's message.
void check(lqcs_status status);

// Global X/Z masks of a full-width Pauli string (leftmost char = qubit n-1),
// validated like Statevector::expectation_value (statevector.cpp:39-60).
struct PauliString {
    uint64_t x = 0, zy = 0;
    uint32_t n_y = 0;
};
PauliString parse_pauli(std::string_view pauli, std::size_t n_qubits);

// Growing op stream; program() views it as an lqcs_program.
class ProgramBuilder {
public:
    // Lowers one unitary instruction with the reference dispatch order
    // (statevector_simulator.cpp:22-85). Returns the op index.
    uint32_t add_instruction(const Instruction& inst);
    // Lowers conj(U) of a unitary instruction onto its qubits + `shift` (the
    // bra side of a density matrix), with the same dispatch.
    uint32_t add_instruction_conj(const Instruction& inst, qubit_t shift);
    // Adds exp(-i theta/2 P) with global masks.
    uint32_t add_pauli_rotation(uint64_t x, uint64_t zy, uint32_t n_y, double theta);
    // Adds a dense 4x4 op (unitary or not) on (q0, q1), local bit 0 <-> q0.
    uint32_t add_two_qubit(qubit_t q0, qubit_t q1, const std::vector<complex_t>& m);
    std::size_t size() const { return ops_.size(); }
    lqcs_program program() const;

private:
    uint32_t push(uint32_t kind, const std::vector<qubit_t>& ctl, const std::vector<qubit_t>& tgt,
                  const std::vector<complex_t>& data);

    std::vector<lqcs_op> ops_;
    std::vector<uint32_t> qubits_;
    std::vector<double> cplx_;
    std::vector<uint64_t> perm_;
};

// Pauli rotation generator of a parametric gate (RX, RY, RZ, PauliRotation)
// on its qubits: the gate equals exp(-i theta/2 P).
PauliString generator_of(const Instruction& inst);

}  // namespace lqcs::detail
