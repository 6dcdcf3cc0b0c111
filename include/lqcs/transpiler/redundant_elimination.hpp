// Redundant-gate elimination (reference include/lqcs/transpiler/redundant_elimination.hpp:9-16):
// removes adjacent inverse pairs (H H, X X, CX CX, S Sdg, RZ(t) RZ(-t), U and
// its inverse, a permutation and its inverse, ...). Two gates are adjacent
// when they act on the same qubit list and no instruction between them
// touches those qubits; a barrier blocks everything. Cascades (X H H X)
// are removed to the fixed point.
#pragma once

#include "pass.hpp"

namespace lqcs {

class RedundantGateElimination final : public Pass {
public:
    QuantumCircuit run(const QuantumCircuit& circuit) const override;
    std::string name() const override { return "redundant_gate_elimination"; }
};

}  // namespace lqcs
