// Redundant-gate elimination. The inverse-pair rules follow the reference's
// src/transpiler/redundant_elimination.cpp:15-55 (structural, no matrix
// products); the removal keeps, per qubit, a stack of the surviving
// instructions that touch it, so that a cancellation exposes the instruction
// beneath at once and cascades resolve in one sweep.
#include "lqcs/transpiler/redundant_elimination.hpp"

#include <vector>

namespace lqcs {

namespace {

bool negated(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == 1 && b.size() == 1 && a[0] == -b[0];
}

// Whether b undoes a (same number of controls; a's kind decides the rule).
bool undoes(const Gate& a, const Gate& b) {
    if (a.n_controls != b.n_controls) return false;
    switch (a.type) {
        case GateType::I: case GateType::X: case GateType::Y: case GateType::Z:
        case GateType::H: case GateType::SWAP:
            return b.type == a.type;  // self-inverse
        case GateType::RX: case GateType::RY: case GateType::RZ: case GateType::P:
        case GateType::RXX: case GateType::RYY: case GateType::RZZ:
            return b.type == a.type && negated(a.params, b.params);
        case GateType::S: return b.type == GateType::Sdg;
        case GateType::Sdg: return b.type == GateType::S;
        case GateType::T: return b.type == GateType::Tdg;
        case GateType::Tdg: return b.type == GateType::T;
        case GateType::U:  // U(t, f, l)^-1 = U(-t, -l, -f)
            return b.type == GateType::U && a.params.size() == 3 && b.params.size() == 3 &&
                   b.params[0] == -a.params[0] && b.params[1] == -a.params[2] && b.params[2] == -a.params[1];
        case GateType::Permutation: {
            if (b.type != GateType::Permutation || a.perm.size() != b.perm.size()) return false;
            for (std::size_t l = 0; l < a.perm.size(); ++l)
                if (b.perm[a.perm[l]] != l) return false;
            return true;
        }
        default:
            return false;  // SX, iSWAP, Pauli rotations, Unitary: left to fusion
    }
}

}  // namespace

QuantumCircuit RedundantGateElimination::run(const QuantumCircuit& circuit) const {
    const auto& in = circuit.instructions();
    std::vector<bool> keep(in.size(), true);
    std::vector<std::vector<std::size_t>> top(circuit.num_qubits());  // surviving instructions per qubit
    for (std::size_t i = 0; i < in.size(); ++i) {
        const Instruction& cur = in[i];
        if (cur.gate.type == GateType::Barrier) {
            for (auto& st : top) st.clear();  // nothing before a barrier can meet anything after it
            continue;
        }
        bool cancelled = false;
        if (cur.gate.is_unitary() && !cur.qubits.empty() && !top[cur.qubits[0]].empty()) {
            const std::size_t j = top[cur.qubits[0]].back();
            bool adjacent = in[j].qubits == cur.qubits;
            for (qubit_t q : cur.qubits) adjacent = adjacent && !top[q].empty() && top[q].back() == j;
            if (adjacent && undoes(in[j].gate, cur.gate)) {
                keep[i] = keep[j] = false;
                for (qubit_t q : cur.qubits) top[q].pop_back();
                cancelled = true;
            }
        }
        if (!cancelled)
            for (qubit_t q : cur.qubits) top[q].push_back(i);
    }
    QuantumCircuit out(circuit.num_qubits(), circuit.num_clbits());
    for (std::size_t i = 0; i < in.size(); ++i)
        if (keep[i]) out.append(in[i]);
    return out;
}

}  // namespace lqcs
