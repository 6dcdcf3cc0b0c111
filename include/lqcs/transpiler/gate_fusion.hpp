// Greedy single-qubit fusion (reference include/lqcs/transpiler/gate_fusion.hpp:11-15):
// runs of uncontrolled one-qubit gates on a qubit become one 2x2 Unitary.
#pragma once

#include "pass.hpp"

namespace lqcs {

class SingleQubitGateFusion final : public Pass {
public:
    QuantumCircuit run(const QuantumCircuit& circuit) const override;
    std::string name() const override { return "single_qubit_gate_fusion"; }
};

}  // namespace lqcs
